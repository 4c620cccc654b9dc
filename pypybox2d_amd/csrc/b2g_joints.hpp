// Joint solver of the island solve (pypybox2d/joints/, called by Island.solve,
// island.py:107-148): each joint type's velocity-constraint setup, velocity solve
// and position solve, on one lane, in FP64 in both pipelines.
//
// A joint's numbers live in its record in HBM (b2g_layout.h B2G_JD_* / B2G_JI_*),
// which stands for the reference joint object's attributes: parameters, accumulated
// impulses and the values the solver computes at setup and reads in later solves or
// later steps.  Values the reference leaves stale between steps therefore stay stale.
// The operation order of every expression follows the reference's (left to right,
// no contraction: the library is built with -ffp-contract=off), and so do its quirks,
// cited at each joint type.
#pragma once

#include "b2g_math.hpp"

namespace b2g {

using JV2 = V2<double>;

struct JV3 {
  double x, y, z;
};
__device__ __forceinline__ JV3 jv3(double x, double y, double z) {
  JV3 r;
  r.x = x;
  r.y = y;
  r.z = z;
  return r;
}
__device__ __forceinline__ JV3 add3(JV3 a, JV3 b) { return jv3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ JV3 sub3(JV3 a, JV3 b) { return jv3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ JV3 neg3(JV3 a) { return jv3(-a.x, -a.y, -a.z); }
__device__ __forceinline__ JV3 scale3(double s, JV3 a) { return jv3(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ double dot3(JV3 a, JV3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ JV3 cross3(JV3 s, JV3 v) {
  return jv3(s.y * v.z - s.z * v.y, s.z * v.x - s.x * v.z, s.x * v.y - s.y * v.x);
}
struct JM33 {
  JV3 c1, c2, c3;
};
// Mat33.solve3x3 (common.py:587-599)
__device__ __forceinline__ JV3 solve33(const JM33& m, JV3 b) {
  double det = dot3(m.c1, cross3(m.c2, m.c3));
  if (det != 0.0) det = 1.0 / det;
  return jv3(det * dot3(b, cross3(m.c2, m.c3)), det * dot3(m.c1, cross3(b, m.c3)),
             det * dot3(m.c1, cross3(m.c2, b)));
}
// Mat33.solve2x2 (common.py:601-614): the upper 2x2 block
__device__ __forceinline__ JV2 solve22(const JM33& m, JV2 v) {
  const double a11 = m.c1.x, a12 = m.c2.x, a21 = m.c1.y, a22 = m.c2.y;
  double det = a11 * a22 - a12 * a21;
  if (det != 0.0) det = 1.0 / det;
  return v2<double>(det * (a22 * v.x - a12 * v.y), det * (a11 * v.y - a21 * v.x));
}
// Mat22.solve (common.py:521-531) and Mat22.inverse (common.py:506-519)
__device__ __forceinline__ JV2 m22_solve(const M22<double>& m, JV2 v) {
  const double a = m.c1x, b = m.c2x, c = m.c1y, d = m.c2y;
  double det = a * d - b * c;
  if (det != 0.0) det = 1.0 / det;
  return v2<double>(det * (d * v.x - b * v.y), det * (a * v.y - c * v.x));
}
__device__ __forceinline__ M22<double> m22_inverse(const M22<double>& m) {
  const double a = m.c1x, b = m.c2x, c = m.c1y, d = m.c2y;
  double det = a * d - b * c;
  if (det != 0.0) det = 1.0 / det;
  M22<double> r;
  r.c1x = det * d;
  r.c1y = -det * c;
  r.c2x = -det * b;
  r.c2y = det * a;
  return r;
}
__device__ __forceinline__ JV2 vdiv(JV2 a, double s) { return v2<double>(a.x / s, a.y / s); }
__device__ __forceinline__ double jlen(JV2 a) { return sqrt(sq(a.x) + sq(a.y)); }
__device__ __forceinline__ double jlen2(JV2 a) { return sq(a.x) + sq(a.y); }
__device__ __forceinline__ Rot<double> jrot(double a) { return rot_of<double>(a); }

// The time step as the joints see it (world.py:53-62).
struct JStep {
  double dt, inv_dt, dt_ratio;
  bool warm;
};

// A joint's bodies in the island: island-local indices (a body outside the island's body list,
// a static one, as -1 - world-local index), inverse masses and inertias, masses, local centres.
// Slots 0..3 are bodies a, b, c, d (c, d: gear joints only).
struct JBodies {
  int idx[4];
  double im[4], iI[4], mass[4];
  JV2 lc[4];
};

// A joint record in HBM.
struct JRec {
  double* d;
  int* i;
  __device__ __forceinline__ JV2 g2(int k) const { return v2<double>(d[k], d[k + 1]); }
  __device__ __forceinline__ void s2(int k, JV2 v) const {
    d[k] = v.x;
    d[k + 1] = v.y;
  }
  __device__ __forceinline__ JV3 g3(int k) const { return jv3(d[k], d[k + 1], d[k + 2]); }
  __device__ __forceinline__ void s3(int k, JV3 v) const {
    d[k] = v.x;
    d[k + 1] = v.y;
    d[k + 2] = v.z;
  }
  __device__ __forceinline__ JM33 g33(int k) const {
    JM33 m;
    m.c1 = g3(k);
    m.c2 = g3(k + 3);
    m.c3 = g3(k + 6);
    return m;
  }
  __device__ __forceinline__ void s33(int k, const JM33& m) const {
    s3(k, m.c1);
    s3(k + 3, m.c2);
    s3(k + 6, m.c3);
  }
  __device__ __forceinline__ M22<double> g22(int k) const {
    M22<double> m;
    m.c1x = d[k];
    m.c1y = d[k + 1];
    m.c2x = d[k + 2];
    m.c2y = d[k + 3];
    return m;
  }
  __device__ __forceinline__ void s22(int k, const M22<double>& m) const {
    d[k] = m.c1x;
    d[k + 1] = m.c1y;
    d[k + 2] = m.c2x;
    d[k + 3] = m.c2y;
  }
  __device__ __forceinline__ bool has(int f) const { return (i[B2G_JI_FLAGS] & f) != 0; }
};

// The two bodies' state of a two-body joint, read from and written back to the island (IO:
// pos(i, c, a), vel(i, v, w), set_pos(i, c, a), set_vel(i, v, w)).
struct JState {
  JV2 ca, cb, va, vb;
  double aa, ab, wa, wb;
};
template <class IO>
__device__ __forceinline__ void jload_vel(const IO& io, const JBodies& B, JState& s) {
  io.vel(B.idx[0], s.va, s.wa);
  io.vel(B.idx[1], s.vb, s.wb);
}
template <class IO>
__device__ __forceinline__ void jload_pos(const IO& io, const JBodies& B, JState& s) {
  io.pos(B.idx[0], s.ca, s.aa);
  io.pos(B.idx[1], s.cb, s.ab);
}
template <class IO>
__device__ __forceinline__ void jstore_vel(const IO& io, const JBodies& B, const JState& s) {
  io.set_vel(B.idx[0], s.va, s.wa);
  io.set_vel(B.idx[1], s.vb, s.wb);
}
template <class IO>
__device__ __forceinline__ void jstore_pos(const IO& io, const JBodies& B, const JState& s) {
  io.set_pos(B.idx[0], s.ca, s.aa);
  io.set_pos(B.idx[1], s.cb, s.ab);
}
__device__ __forceinline__ JV2 lever(double angle, JV2 local_anchor, JV2 local_center) {
  return rmul(jrot(angle), local_anchor - local_center);
}
__device__ __forceinline__ void zero_limit_impulse(const JRec& J) {
  J.d[B2G_JD_IMP_Z] = J.d[B2G_JD_IMP_Z] - J.d[B2G_JD_IMP_Z];  // impulse -= (0, 0, impulse.z)
}

constexpr double kJLinearSlop = 0.005;                                     // settings.py:61
constexpr double kJAngularSlop = 2.0 / 180.0 * 3.141592653589793;           // settings.py:65
constexpr double kJMaxLinearCorrection = 0.2;                              // settings.py:90
constexpr double kJMaxAngularCorrection = 8.0 / 180.0 * 3.141592653589793;  // settings.py:94
constexpr double kJPi = 3.141592653589793;
constexpr double kJEpsilon = 2.220446049250313080847e-16;

// ---- revolute (revolute.py:214-453) ----
template <class IO>
__device__ void revolute_init(const JRec& J, const JBodies& B, const IO& io, const JStep& st) {
  const double ma = B.im[0], mb = B.im[1], ia = B.iI[0], ib = B.iI[1];
  JState s;
  jload_pos(io, B, s);
  jload_vel(io, B, s);
  const JV2 ra = lever(s.aa, J.g2(B2G_JD_LA_X), B.lc[0]);
  const JV2 rb = lever(s.ab, J.g2(B2G_JD_LB_X), B.lc[1]);
  J.s2(B2G_JD_RA_X, ra);
  J.s2(B2G_JD_RB_X, rb);
  JM33 m;
  m.c3 = jv3(-ra.y * ia - rb.y * ib, ra.x * ia + rb.x * ib, ia + ib);
  m.c2 = jv3(-ra.y * ra.x * ia - rb.y * rb.x * ib, ma + mb + ra.x * ra.x * ia + rb.x * rb.x * ib, m.c3.y);
  m.c1 = jv3(ma + mb + ra.y * ra.y * ia + rb.y * rb.y * ib, m.c2.x, m.c3.x);
  J.s33(B2G_JD_REV_MASS, m);
  double motor_mass = ia + ib;
  const bool fixed_rotation = motor_mass == 0.0;
  if (motor_mass > 0.0) motor_mass = 1.0 / motor_mass;
  J.d[B2G_JD_REV_MOTOR_MASS] = motor_mass;
  if (!J.has(B2G_JF_MOTOR) || fixed_rotation) J.d[B2G_JD_MOTOR_IMP] = 0.0;
  int& state = J.i[B2G_JI_STATE];
  if (J.has(B2G_JF_LIMIT) && !fixed_rotation) {
    const double joint_angle = s.ab - s.aa - J.d[B2G_JD_REF];
    const double lower = J.d[B2G_JD_LOWER], upper = J.d[B2G_JD_UPPER];
    if (fabs(upper - lower) < 2.0 * kJAngularSlop) {
      state = B2G_LIMIT_EQUAL;
    } else if (joint_angle <= lower) {
      if (state != B2G_LIMIT_AT_LOWER) zero_limit_impulse(J);
      state = B2G_LIMIT_AT_LOWER;
    } else if (joint_angle >= upper) {
      if (state != B2G_LIMIT_AT_UPPER) zero_limit_impulse(J);
      state = B2G_LIMIT_AT_UPPER;
    } else {
      state = B2G_LIMIT_INACTIVE;
      zero_limit_impulse(J);
    }
  } else {
    state = B2G_LIMIT_INACTIVE;
  }
  if (st.warm) {
    const JV3 imp = scale3(st.dt_ratio, J.g3(B2G_JD_IMP_X));
    J.s3(B2G_JD_IMP_X, imp);
    const double mi = J.d[B2G_JD_MOTOR_IMP] * st.dt_ratio;
    J.d[B2G_JD_MOTOR_IMP] = mi;
    const JV2 P = v2<double>(imp.x, imp.y);
    s.va = s.va - scale(ma, P);
    s.wa -= ia * (cross(ra, P) + mi + imp.z);
    s.vb = s.vb + scale(mb, P);
    s.wb += ib * (cross(rb, P) + mi + imp.z);
  } else {
    J.s3(B2G_JD_IMP_X, jv3(0.0, 0.0, 0.0));
    J.d[B2G_JD_MOTOR_IMP] = 0.0;
  }
  jstore_vel(io, B, s);
}

template <class IO>
__device__ void revolute_solve_velocity(const JRec& J, const JBodies& B, const IO& io, const JStep& st) {
  const double ma = B.im[0], mb = B.im[1], ia = B.iI[0], ib = B.iI[1];
  JState s;
  jload_vel(io, B, s);
  const JV2 ra = J.g2(B2G_JD_RA_X), rb = J.g2(B2G_JD_RB_X);
  const JM33 mass = J.g33(B2G_JD_REV_MASS);
  const bool fixed_rotation = (ia + ib) == 0.0;
  const int state = J.i[B2G_JI_STATE];
  if (J.has(B2G_JF_MOTOR) && state != B2G_LIMIT_EQUAL && !fixed_rotation) {
    const double Cdot = s.wb - s.wa - J.d[B2G_JD_MOTOR_SPEED];
    double impulse = J.d[B2G_JD_REV_MOTOR_MASS] * (-Cdot);
    const double old = J.d[B2G_JD_MOTOR_IMP];
    const double max_impulse = st.dt * J.d[B2G_JD_MAX_MOTOR];
    J.d[B2G_JD_MOTOR_IMP] = pclamp(old + impulse, -max_impulse, max_impulse);
    impulse = J.d[B2G_JD_MOTOR_IMP] - old;
    s.wa -= ia * impulse;
    s.wb += ib * impulse;
  }
  JV3 acc = J.g3(B2G_JD_IMP_X);
  if (J.has(B2G_JF_LIMIT) && state != B2G_LIMIT_INACTIVE && !fixed_rotation) {
    const JV2 Cdot1 = s.vb + cross_sv(s.wb, rb) - s.va - cross_sv(s.wa, ra);
    const double Cdot2 = s.wb - s.wa;
    JV3 impulse = neg3(solve33(mass, jv3(Cdot1.x, Cdot1.y, Cdot2)));
    if (state == B2G_LIMIT_EQUAL) {
      acc = add3(acc, impulse);
    } else {
      const double new_impulse = acc.z + impulse.z;
      const bool reduce = state == B2G_LIMIT_AT_LOWER ? new_impulse < 0.0 : new_impulse > 0.0;
      if (reduce) {
        const JV2 rhs = -Cdot1 + scale(acc.z, v2<double>(mass.c3.x, mass.c3.y));
        const JV2 reduced = solve22(mass, rhs);
        impulse = jv3(reduced.x, reduced.y, -acc.z);
        acc = add3(acc, jv3(reduced.x, reduced.y, -acc.z));
      } else {
        acc = add3(acc, impulse);
      }
    }
    const JV2 P = v2<double>(impulse.x, impulse.y);
    s.va = s.va - scale(ma, P);
    s.wa -= ia * (cross(ra, P) + impulse.z);
    s.vb = s.vb + scale(mb, P);
    s.wb += ib * (cross(rb, P) + impulse.z);
  } else {
    const JV2 Cdot = s.vb + cross_sv(s.wb, rb) - s.va - cross_sv(s.wa, ra);
    const JV2 impulse = solve22(mass, -Cdot);
    acc = add3(acc, jv3(impulse.x, impulse.y, 0.0));
    s.va = s.va - scale(ma, impulse);
    s.wa -= ia * cross(ra, impulse);
    s.vb = s.vb + scale(mb, impulse);
    s.wb += ib * cross(rb, impulse);
  }
  J.s3(B2G_JD_IMP_X, acc);
  jstore_vel(io, B, s);
}

template <class IO>
__device__ bool revolute_solve_position(const JRec& J, const JBodies& B, const IO& io) {
  const double ma = B.im[0], mb = B.im[1], ia = B.iI[0], ib = B.iI[1];
  JState s;
  jload_pos(io, B, s);
  const bool fixed_rotation = (ia + ib) == 0.0;
  double angular_error = 0.0;
  const int state = J.i[B2G_JI_STATE];
  if (J.has(B2G_JF_LIMIT) && state != B2G_LIMIT_INACTIVE && !fixed_rotation) {
    const double angle = s.ab - s.aa - J.d[B2G_JD_REF];
    const double motor_mass = J.d[B2G_JD_REV_MOTOR_MASS];
    double limit_impulse = 0.0, C;
    if (state == B2G_LIMIT_EQUAL) {
      C = pclamp(angle - J.d[B2G_JD_LOWER], -kJMaxAngularCorrection, kJMaxAngularCorrection);
      limit_impulse = -motor_mass * C;
      angular_error = fabs(C);
    } else if (state == B2G_LIMIT_AT_LOWER) {
      C = angle - J.d[B2G_JD_LOWER];
      angular_error = -C;
      C = pclamp(C + kJAngularSlop, -kJMaxAngularCorrection, 0.0);
      limit_impulse = -motor_mass * C;
    } else {
      C = angle - J.d[B2G_JD_UPPER];
      angular_error = C;
      C = pclamp(C - kJAngularSlop, 0.0, kJMaxAngularCorrection);
      limit_impulse = -motor_mass * C;
    }
    s.aa -= ia * limit_impulse;
    s.ab += ib * limit_impulse;
  }
  const JV2 ra = lever(s.aa, J.g2(B2G_JD_LA_X), B.lc[0]);
  const JV2 rb = lever(s.ab, J.g2(B2G_JD_LB_X), B.lc[1]);
  const JV2 C = s.cb + rb - s.ca - ra;
  const double position_error = jlen(C);
  const double rax = ra.x, ray = ra.y, rbx = rb.x, rby = rb.y;
  M22<double> K;
  K.c1x = ma + mb + ia * ray * ray + ib * rby * rby;
  K.c1y = -ia * rax * ray - ib * rbx * rby;
  K.c2x = K.c1y;
  K.c2y = ma + mb + ia * rax * rax + ib * rbx * rbx;
  const JV2 impulse = -m22_solve(K, C);
  s.ca = s.ca - scale(ma, impulse);
  s.aa -= ia * cross(ra, impulse);
  s.cb = s.cb + scale(mb, impulse);
  s.ab += ib * cross(rb, impulse);
  jstore_pos(io, B, s);
  return position_error <= kJLinearSlop && angular_error <= kJAngularSlop;
}

// ---- prismatic (prismatic.py:271-513) ----
__device__ __forceinline__ JM33 prismatic_k(double ma, double mb, double ia, double ib, double s1, double s2,
                                            double a1, double a2) {
  const double k11 = ma + mb + ia * sq(s1) + ib * sq(s2);
  const double k12 = ia * s1 + ib * s2;
  const double k13 = ia * s1 * a1 + ib * s2 * a2;
  double k22 = ia + ib;
  if (k22 == 0.0) k22 = 1.0;
  const double k23 = ia * a1 + ib * a2;
  const double k33 = ma + mb + ia * sq(a1) + ib * sq(a2);
  JM33 k;
  k.c1 = jv3(k11, k12, k13);
  k.c2 = jv3(k12, k22, k23);
  k.c3 = jv3(k13, k23, k33);
  return k;
}

template <class IO>
__device__ void prismatic_init(const JRec& J, const JBodies& B, const IO& io, const JStep& st) {
  const double ma = B.im[0], mb = B.im[1], ia = B.iI[0], ib = B.iI[1];
  JState s;
  jload_pos(io, B, s);
  jload_vel(io, B, s);
  const Rot<double> qa = jrot(s.aa);
  const JV2 ra = rmul(qa, J.g2(B2G_JD_LA_X) - B.lc[0]);
  const JV2 rb = lever(s.ab, J.g2(B2G_JD_LB_X), B.lc[1]);
  J.s2(B2G_JD_RA_X, ra);
  J.s2(B2G_JD_RB_X, rb);
  const JV2 d = s.cb + rb - s.ca - ra;
  const JV2 axis = rmul(qa, J.g2(B2G_JD_AXIS_X));
  J.s2(B2G_JD_PRI_AXIS_X, axis);
  const double a1 = cross(d + ra, axis), a2 = cross(rb, axis);
  J.d[B2G_JD_PRI_A1] = a1;
  J.d[B2G_JD_PRI_A2] = a2;
  double motor_mass = ma + mb + ia * sq(a1) + ib * sq(a2);
  if (motor_mass > 0.0) motor_mass = 1.0 / motor_mass;
  J.d[B2G_JD_PRI_MOTOR_MASS] = motor_mass;
  const JV2 perp = rmul(qa, J.g2(B2G_JD_PERP_X));
  J.s2(B2G_JD_PRI_PERP_X, perp);
  const double s1 = cross(d + ra, perp), s2 = cross(rb, perp);
  J.d[B2G_JD_PRI_S1] = s1;
  J.d[B2G_JD_PRI_S2] = s2;
  J.s33(B2G_JD_PRI_K, prismatic_k(ma, mb, ia, ib, s1, s2, a1, a2));
  int& state = J.i[B2G_JI_STATE];
  if (J.has(B2G_JF_LIMIT)) {
    const double jt = dot(axis, d);
    const double lower = J.d[B2G_JD_LOWER], upper = J.d[B2G_JD_UPPER];
    if (fabs(upper - lower) < 2.0 * kJLinearSlop) {
      state = B2G_LIMIT_EQUAL;
    } else if (jt <= lower) {
      if (state != B2G_LIMIT_AT_LOWER) {
        state = B2G_LIMIT_AT_LOWER;
        zero_limit_impulse(J);
      }
    } else if (jt >= upper) {
      if (state != B2G_LIMIT_AT_UPPER) {
        state = B2G_LIMIT_AT_UPPER;
        zero_limit_impulse(J);
      }
    } else {
      state = B2G_LIMIT_INACTIVE;
      zero_limit_impulse(J);
    }
  } else {
    state = B2G_LIMIT_INACTIVE;
    zero_limit_impulse(J);
  }
  if (!J.has(B2G_JF_MOTOR)) J.d[B2G_JD_MOTOR_IMP] = 0.0;
  if (st.warm) {
    const JV3 imp = scale3(st.dt_ratio, J.g3(B2G_JD_IMP_X));
    J.s3(B2G_JD_IMP_X, imp);
    const double mi = J.d[B2G_JD_MOTOR_IMP] * st.dt_ratio;
    J.d[B2G_JD_MOTOR_IMP] = mi;
    const JV2 P = scale(imp.x, perp) + scale(mi + imp.z, axis);
    const double L1 = imp.x * s1 + imp.y + (mi + imp.z) * a1;
    const double L2 = imp.x * s2 + imp.y + (mi + imp.z) * a2;
    s.va = s.va - scale(ma, P);
    s.wa -= ia * L1;
    s.vb = s.vb + scale(mb, P);
    s.wb += ib * L2;
  } else {
    J.s3(B2G_JD_IMP_X, jv3(0.0, 0.0, 0.0));
    J.d[B2G_JD_MOTOR_IMP] = 0.0;
  }
  jstore_vel(io, B, s);
}

template <class IO>
__device__ void prismatic_solve_velocity(const JRec& J, const JBodies& B, const IO& io, const JStep& st) {
  const double ma = B.im[0], mb = B.im[1], ia = B.iI[0], ib = B.iI[1];
  JState s;
  jload_vel(io, B, s);
  const double a1 = J.d[B2G_JD_PRI_A1], a2 = J.d[B2G_JD_PRI_A2];
  const double s1 = J.d[B2G_JD_PRI_S1], s2 = J.d[B2G_JD_PRI_S2];
  const JV2 perp = J.g2(B2G_JD_PRI_PERP_X), axis = J.g2(B2G_JD_PRI_AXIS_X);
  const JM33 k = J.g33(B2G_JD_PRI_K);
  const int state = J.i[B2G_JI_STATE];
  if (J.has(B2G_JF_MOTOR) && state != B2G_LIMIT_EQUAL) {
    const double Cdot = dot(axis, s.vb - s.va) + a2 * s.wb - a1 * s.wa;
    double impulse = J.d[B2G_JD_PRI_MOTOR_MASS] * (J.d[B2G_JD_MOTOR_SPEED] - Cdot);
    const double old = J.d[B2G_JD_MOTOR_IMP];
    const double max_impulse = st.dt * J.d[B2G_JD_MAX_MOTOR];
    J.d[B2G_JD_MOTOR_IMP] = pclamp(old + impulse, -max_impulse, max_impulse);
    impulse = J.d[B2G_JD_MOTOR_IMP] - old;
    const JV2 P = scale(impulse, axis);
    const double L1 = impulse * a1, L2 = impulse * a2;
    s.va = s.va - scale(ma, P);
    s.wa -= ia * L1;
    s.vb = s.vb + scale(mb, P);
    s.wb += ib * L2;
  }
  const JV2 Cdot1 = v2<double>(dot(perp, s.vb - s.va) + s2 * s.wb - s1 * s.wa, s.wb - s.wa);
  JV2 P;
  double L1, L2;
  JV3 acc = J.g3(B2G_JD_IMP_X);
  if (J.has(B2G_JF_LIMIT) && state != B2G_LIMIT_INACTIVE) {
    const double Cdot2 = dot(axis, s.vb - s.va) + a2 * s.wb - a1 * s.wa;
    const JV3 old = acc;
    JV3 df = solve33(k, neg3(jv3(Cdot1.x, Cdot1.y, Cdot2)));
    acc = add3(acc, df);
    double iz;
    if (state == B2G_LIMIT_AT_LOWER) iz = pmax(acc.z, 0.0);
    else if (state == B2G_LIMIT_AT_UPPER) iz = pmin(acc.z, 0.0);
    else iz = acc.z;
    const JV2 bb = -Cdot1 - scale(acc.z - old.z, v2<double>(k.c3.x, k.c3.y));
    const JV2 f2r = solve22(k, bb) + v2<double>(old.x, old.y);
    acc = jv3(f2r.x, f2r.y, iz);
    df = sub3(acc, old);
    P = scale(df.x, perp) + scale(df.z, axis);
    L1 = df.x * s1 + df.y + df.z * a1;
    L2 = df.x * s2 + df.y + df.z * a2;
  } else {
    const JV2 df = solve22(k, -Cdot1);
    acc = add3(acc, jv3(df.x, df.y, 0.0));
    P = scale(df.x, perp);
    L1 = df.x * s1 + df.y;
    L2 = df.x * s2 + df.y;
  }
  J.s3(B2G_JD_IMP_X, acc);
  s.va = s.va - scale(ma, P);
  s.wa -= ia * L1;
  s.vb = s.vb + scale(mb, P);
  s.wb += ib * L2;
  jstore_vel(io, B, s);
}

template <class IO>
__device__ bool prismatic_solve_position(const JRec& J, const JBodies& B, const IO& io) {
  const double ma = B.im[0], mb = B.im[1], ia = B.iI[0], ib = B.iI[1];
  JState s;
  jload_pos(io, B, s);
  const Rot<double> qa = jrot(s.aa);
  const JV2 ra = rmul(qa, J.g2(B2G_JD_LA_X) - B.lc[0]);
  const JV2 rb = lever(s.ab, J.g2(B2G_JD_LB_X), B.lc[1]);
  const JV2 d = s.cb + rb - s.ca - ra;
  const JV2 axis = rmul(qa, J.g2(B2G_JD_AXIS_X));
  const double a1 = cross(d + ra, axis), a2 = cross(rb, axis);
  const JV2 perp = rmul(qa, J.g2(B2G_JD_PERP_X));
  const double s1 = cross(d + ra, perp), s2 = cross(rb, perp);
  const JV2 C1 = v2<double>(dot(perp, d), s.ab - s.aa - J.d[B2G_JD_REF]);
  double linear_error = fabs(C1.x);
  const double angular_error = fabs(C1.y);
  bool active = false;
  double C2 = 0.0;
  if (J.has(B2G_JF_LIMIT)) {
    const double translation = dot(axis, d);
    const double lower = J.d[B2G_JD_LOWER], upper = J.d[B2G_JD_UPPER];
    if (fabs(upper - lower) < 2.0 * kJLinearSlop) {
      C2 = pclamp(translation, -kJMaxLinearCorrection, kJMaxLinearCorrection);
      linear_error = pmax(linear_error, translation);  // the reference takes no abs here
      active = true;
    } else if (translation <= lower) {
      C2 = pclamp(translation - lower + kJLinearSlop, -kJMaxLinearCorrection, 0.0);
      linear_error = pmax(linear_error, lower - translation);
      active = true;
    } else if (translation >= upper) {
      C2 = pclamp(translation - upper - kJLinearSlop, 0.0, kJMaxLinearCorrection);
      linear_error = pmax(linear_error, translation - upper);
      active = true;
    }
  }
  JV3 impulse;
  if (active) {
    const JM33 k = prismatic_k(ma, mb, ia, ib, s1, s2, a1, a2);
    J.s33(B2G_JD_PRI_K, k);
    impulse = solve33(k, neg3(jv3(C1.x, C1.y, C2)));
  } else {
    const double k11 = ma + mb + ia * s1 * s1 + ib * s2 * s2;
    const double k12 = ia * s1 + ib * s2;
    double k22 = ia + ib;
    if (k22 == 0.0) k22 = 1.0;
    JM33 k = J.g33(B2G_JD_PRI_K);  // the setup's third column stays
    k.c1 = jv3(k11, k12, 0.0);
    k.c2 = jv3(k12, k22, 0.0);
    J.s33(B2G_JD_PRI_K, k);
    const JV2 i1 = solve22(k, -C1);
    impulse = jv3(i1.x, i1.y, 0.0);
  }
  const JV2 P = scale(impulse.x, perp) + scale(impulse.z, axis);
  const double L1 = impulse.x * s1 + impulse.y + impulse.z * a1;
  const double L2 = impulse.x * s2 + impulse.y + impulse.z * a2;
  s.ca = s.ca - scale(ma, P);
  s.cb = s.cb + scale(mb, P);
  s.aa -= ia * L1;
  s.ab += ib * L2;
  jstore_pos(io, B, s);
  return linear_error <= kJLinearSlop && angular_error <= kJAngularSlop;
}

// ---- distance (distance.py:170-311): the effective mass is taken with the separation before
// normalisation (`self._u *= 1 / length` rebinds the attribute; the local u keeps the old
// vector), and the solves use the normalised attribute ----
template <class IO>
__device__ void distance_init(const JRec& J, const JBodies& B, const IO& io, const JStep& st) {
  const double ma = B.im[0], mb = B.im[1], ia = B.iI[0], ib = B.iI[1];
  JState s;
  jload_pos(io, B, s);
  jload_vel(io, B, s);
  const JV2 ra = lever(s.aa, J.g2(B2G_JD_LA_X), B.lc[0]);
  const JV2 rb = lever(s.ab, J.g2(B2G_JD_LB_X), B.lc[1]);
  J.s2(B2G_JD_RA_X, ra);
  J.s2(B2G_JD_RB_X, rb);
  const JV2 u = s.cb + rb - s.ca - ra;
  const double length = jlen(u);
  const JV2 un = length > kJLinearSlop ? scale(1.0 / length, u) : v2<double>(0.0, 0.0);
  J.s2(B2G_JD_DIS_U_X, un);
  const double cr_au = sq(cross(ra, u)), cr_bu = sq(cross(rb, u));
  const double inv_mass = ma + ia * cr_au + mb + ib * cr_bu;
  const double mass = inv_mass != 0.0 ? 1.0 / inv_mass : 0.0;
  J.d[B2G_JD_DIS_MASS] = mass;
  if (J.d[B2G_JD_FREQ] > 0.0) {
    const double C = length - J.d[B2G_JD_LENGTH];
    const double omega = 2.0 * kJPi * J.d[B2G_JD_FREQ];
    const double dd = 2.0 * mass * J.d[B2G_JD_DAMP] * omega;
    const double k = mass * sq(omega);
    const double dt = st.dt;
    double gamma = dt * (dd + dt * k);
    gamma = gamma != 0.0 ? 1.0 / gamma : 0.0;
    J.d[B2G_JD_DIS_GAMMA] = gamma;
    J.d[B2G_JD_DIS_BIAS] = C * dt * k * gamma;
    const double m2 = inv_mass + gamma;
    J.d[B2G_JD_DIS_MASS] = m2 != 0.0 ? 1.0 / m2 : 0.0;
  }
  if (st.warm) {
    const double imp = J.d[B2G_JD_IMP_X] * st.dt_ratio;
    J.d[B2G_JD_IMP_X] = imp;
    const JV2 P = scale(imp, un);
    s.va = s.va - scale(ma, P);
    s.wa -= ia * cross(ra, P);
    s.vb = s.vb + scale(mb, P);
    s.wb += ib * cross(rb, P);
  } else {
    J.d[B2G_JD_IMP_X] = 0.0;
  }
  jstore_vel(io, B, s);
}

template <class IO>
__device__ void distance_solve_velocity(const JRec& J, const JBodies& B, const IO& io) {
  JState s;
  jload_vel(io, B, s);
  const JV2 ra = J.g2(B2G_JD_RA_X), rb = J.g2(B2G_JD_RB_X), u = J.g2(B2G_JD_DIS_U_X);
  const JV2 vpa = s.va + cross_sv(s.wa, ra), vpb = s.vb + cross_sv(s.wb, rb);
  const double Cdot = dot(u, vpb - vpa);
  const double impulse =
      -J.d[B2G_JD_DIS_MASS] * (Cdot + J.d[B2G_JD_DIS_BIAS] + J.d[B2G_JD_DIS_GAMMA] * J.d[B2G_JD_IMP_X]);
  J.d[B2G_JD_IMP_X] += impulse;
  const JV2 P = scale(impulse, u);
  s.va = s.va - scale(B.im[0], P);
  s.wa -= B.iI[0] * cross(ra, P);
  s.vb = s.vb + scale(B.im[1], P);
  s.wb += B.iI[1] * cross(rb, P);
  jstore_vel(io, B, s);
}

template <class IO>
__device__ bool distance_solve_position(const JRec& J, const JBodies& B, const IO& io) {
  if (J.d[B2G_JD_FREQ] > 0.0) return true;  // no position correction for soft constraints
  JState s;
  jload_pos(io, B, s);
  const JV2 ra = lever(s.aa, J.g2(B2G_JD_LA_X), B.lc[0]);
  const JV2 rb = lever(s.ab, J.g2(B2G_JD_LB_X), B.lc[1]);
  JV2 u = s.cb + rb - s.ca - ra;
  const double length = jlen(u);
  u = normalized(u);
  const double C = length - J.d[B2G_JD_LENGTH];
  const double impulse = -J.d[B2G_JD_DIS_MASS] * pclamp(C, -kJMaxLinearCorrection, kJMaxLinearCorrection);
  const JV2 P = scale(impulse, u);
  s.ca = s.ca - scale(B.im[0], P);
  s.aa -= B.iI[0] * cross(ra, P);
  s.cb = s.cb + scale(B.im[1], P);
  s.ab += B.iI[1] * cross(rb, P);
  jstore_pos(io, B, s);
  return fabs(C) < kJLinearSlop;
}

// ---- pulley (pulley.py:181-320): the warm start and the position correction move the bodies
// by the effective masses (the local ma, mb are rebound before they are used) ----
template <class IO>
__device__ void pulley_init(const JRec& J, const JBodies& B, const IO& io, const JStep& st) {
  double ma = B.im[0], mb = B.im[1];
  const double ia = B.iI[0], ib = B.iI[1];
  JState s;
  jload_pos(io, B, s);
  jload_vel(io, B, s);
  const JV2 ra = lever(s.aa, J.g2(B2G_JD_LA_X), B.lc[0]);
  const JV2 rb = lever(s.ab, J.g2(B2G_JD_LB_X), B.lc[1]);
  J.s2(B2G_JD_RA_X, ra);
  J.s2(B2G_JD_RB_X, rb);
  JV2 ua = s.ca + ra - J.g2(B2G_JD_G0_X);
  JV2 ub = s.cb + rb - J.g2(B2G_JD_G1_X);
  const double length_a = jlen(ua), length_b = jlen(ub);
  ua = length_a > 10.0 * kJLinearSlop ? vdiv(ua, length_a) : v2<double>(0.0, 0.0);
  ub = length_b > 10.0 * kJLinearSlop ? vdiv(ub, length_b) : v2<double>(0.0, 0.0);
  J.s2(B2G_JD_PUL_UA_X, ua);
  J.s2(B2G_JD_PUL_UB_X, ub);
  const double rua = cross(ra, ua), rub = cross(rb, ub);
  ma = ma + ia * sq(rua);
  mb = mb + ib * sq(rub);
  const double ratio = J.d[B2G_JD_RATIO];
  double mass = ma + ratio * ratio * mb;
  if (mass > 0.0) mass = 1.0 / mass;
  J.d[B2G_JD_PUL_MASS] = mass;
  if (st.warm) {
    const double imp = J.d[B2G_JD_IMP_X] * st.dt_ratio;
    J.d[B2G_JD_IMP_X] = imp;
    const JV2 Pa = scale(-imp, ua);
    const JV2 Pb = scale(-ratio * imp, ub);
    s.va = s.va + scale(ma, Pa);
    s.wa += ia * cross(ra, Pa);
    s.vb = s.vb + scale(mb, Pb);
    s.wb += ib * cross(rb, Pb);
  } else {
    J.d[B2G_JD_IMP_X] = 0.0;
  }
  jstore_vel(io, B, s);
}

template <class IO>
__device__ void pulley_solve_velocity(const JRec& J, const JBodies& B, const IO& io) {
  JState s;
  jload_vel(io, B, s);
  const JV2 ra = J.g2(B2G_JD_RA_X), rb = J.g2(B2G_JD_RB_X);
  const JV2 ua = J.g2(B2G_JD_PUL_UA_X), ub = J.g2(B2G_JD_PUL_UB_X);
  const double ratio = J.d[B2G_JD_RATIO];
  const JV2 vpa = s.va + cross_sv(s.wa, ra), vpb = s.vb + cross_sv(s.wb, rb);
  const double Cdot = -dot(ua, vpa) - ratio * dot(ub, vpb);
  const double impulse = -J.d[B2G_JD_PUL_MASS] * Cdot;
  J.d[B2G_JD_IMP_X] += impulse;
  const JV2 Pa = scale(-impulse, ua);
  const JV2 Pb = scale(-ratio * impulse, ub);
  s.va = s.va + scale(B.im[0], Pa);
  s.wa += B.iI[0] * cross(ra, Pa);
  s.vb = s.vb + scale(B.im[1], Pb);
  s.wb += B.iI[1] * cross(rb, Pb);
  jstore_vel(io, B, s);
}

template <class IO>
__device__ bool pulley_solve_position(const JRec& J, const JBodies& B, const IO& io) {
  double ma = B.im[0], mb = B.im[1];
  const double ia = B.iI[0], ib = B.iI[1];
  JState s;
  jload_pos(io, B, s);
  const JV2 ra = lever(s.aa, J.g2(B2G_JD_LA_X), B.lc[0]);
  const JV2 rb = lever(s.ab, J.g2(B2G_JD_LB_X), B.lc[1]);
  JV2 ua = s.ca + ra - J.g2(B2G_JD_G0_X);
  JV2 ub = s.cb + rb - J.g2(B2G_JD_G1_X);
  const double length_a = jlen(ua), length_b = jlen(ub);
  ua = length_a > 10.0 * kJLinearSlop ? vdiv(ua, length_a) : v2<double>(0.0, 0.0);
  ub = length_b > 10.0 * kJLinearSlop ? vdiv(ub, length_b) : v2<double>(0.0, 0.0);
  const double rua = cross(ra, ua), rub = cross(rb, ub);
  ma = ma + ia * sq(rua);
  mb = mb + ib * sq(rub);
  const double ratio = J.d[B2G_JD_RATIO];
  double mass = ma + sq(ratio) * mb;
  if (mass > 0.0) mass = 1.0 / mass;
  const double C = J.d[B2G_JD_LENGTH] - length_a - ratio * length_b;
  const double linear_error = fabs(C);
  const double impulse = -mass * C;
  const JV2 Pa = scale(-impulse, ua);
  const JV2 Pb = scale(-ratio * impulse, ub);
  s.ca = s.ca + scale(ma, Pa);
  s.aa += ia * cross(ra, Pa);
  s.cb = s.cb + scale(mb, Pb);
  s.ab += ib * cross(rb, Pb);
  jstore_pos(io, B, s);
  return linear_error < kJLinearSlop;
}

// ---- mouse (mouse.py:150-229): one body, body_b (slot 1) ----
template <class IO>
__device__ bool mouse_init(const JRec& J, const JBodies& B, const IO& io, const JStep& st) {
  const int ix = B.idx[1];
  JV2 cb, vb;
  double ab, wb;
  io.pos(ix, cb, ab);
  io.vel(ix, vb, wb);
  const double mb = B.im[1], ib = B.iI[1];
  const double mass = B.mass[1];
  const double omega = 2.0 * kJPi * J.d[B2G_JD_FREQ];
  const double d = 2.0 * mass * J.d[B2G_JD_DAMP] * omega;
  const double k = mass * sq(omega);
  const double dt = st.dt;
  const bool ok = d + dt * k > kJEpsilon;  // the reference asserts this (mouse.py:173)
  double gamma = dt * (d + dt * k);
  if (gamma != 0.0) gamma = 1.0 / gamma;
  const double beta = dt * k * gamma;
  J.d[B2G_JD_MOU_GAMMA] = gamma;
  J.d[B2G_JD_MOU_BETA] = beta;
  const JV2 rb = lever(ab, J.g2(B2G_JD_LB_X), B.lc[1]);
  J.s2(B2G_JD_RB_X, rb);
  M22<double> K;
  K.c1x = mb + ib * sq(rb.y) + gamma;
  K.c1y = -ib * rb.x * rb.y;
  K.c2x = K.c1y;
  K.c2y = mb + ib * sq(rb.x) + gamma;
  J.s22(B2G_JD_MOU_MASS, m22_inverse(K));
  J.s2(B2G_JD_MOU_C_X, scale(beta, cb + rb - J.g2(B2G_JD_G0_X)));
  wb *= 0.98;
  if (st.warm) {
    const JV2 imp = scale(st.dt_ratio, J.g2(B2G_JD_IMP_X));
    J.s2(B2G_JD_IMP_X, imp);
    vb = vb + scale(mb, imp);
    wb += ib * cross(rb, imp);
  } else {
    J.s2(B2G_JD_IMP_X, v2<double>(0.0, 0.0));
  }
  io.set_vel(ix, vb, wb);
  return ok;
}

template <class IO>
__device__ void mouse_solve_velocity(const JRec& J, const JBodies& B, const IO& io, const JStep& st) {
  const int ix = B.idx[1];
  JV2 vb;
  double wb;
  io.vel(ix, vb, wb);
  const JV2 rb = J.g2(B2G_JD_RB_X);
  const JV2 Cdot = vb + cross_sv(wb, rb);
  JV2 acc = J.g2(B2G_JD_IMP_X);
  JV2 impulse = mmul(J.g22(B2G_JD_MOU_MASS), -(Cdot + J.g2(B2G_JD_MOU_C_X) + scale(J.d[B2G_JD_MOU_GAMMA], acc)));
  const JV2 old = acc;
  acc = acc + impulse;
  const double max_impulse = st.dt * J.d[B2G_JD_MAX_MOTOR];
  if (jlen2(acc) > sq(max_impulse)) acc = scale(max_impulse / jlen(acc), acc);
  J.s2(B2G_JD_IMP_X, acc);
  impulse = acc - old;
  vb = vb + scale(B.im[1], impulse);
  wb += B.iI[1] * cross(rb, impulse);
  io.set_vel(ix, vb, wb);
}

// ---- gear (gear.py:262-388).  Its constructor takes bodies b and d from joint_a as well
// (gear.py:91-92), so a == b and c == d.  The velocity solve keeps the linear velocities it
// captured at setup (only the angular ones are read again, gear.py:337-340), and the position
// solve starts from the positions captured at setup (gear.py:356-358) and writes them back with
// its correction. ----
template <class IO>
__device__ void gear_init(const JRec& J, const JBodies& B, const IO& io, const JStep& st) {
  JV2 Cs[4], Vs[4];
  double As[4], Ws[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    io.pos(B.idx[k], Cs[k], As[k]);
    io.vel(B.idx[k], Vs[k], Ws[k]);
  }
  const Rot<double> qa = jrot(As[0]), qb = jrot(As[1]), qc = jrot(As[2]), qd = jrot(As[3]);
  const double ia = B.iI[0], ib = B.iI[1], ic = B.iI[2], id_ = B.iI[3];
  const double ma = B.im[0], mb = B.im[1], mc = B.im[2], md = B.im[3];
  const double ratio = J.d[B2G_JD_RATIO];
  double mass = 0.0;
  JV2 j_vac, j_vbd;
  double j_wa, j_wb, j_wc, j_wd;
  if (J.has(B2G_JF_REVOLUTE_A)) {
    j_vac = v2<double>(0.0, 0.0);
    j_wa = 1.0;
    j_wc = 1.0;
    mass += ia + ic;
  } else {
    const JV2 u = rmul(qc, J.g2(B2G_JD_AXIS_X));
    const JV2 rc = rmul(qc, J.g2(B2G_JD_G0_X) - B.lc[2]);
    const JV2 ra = rmul(qa, J.g2(B2G_JD_LA_X) - B.lc[0]);
    j_vac = u;
    j_wc = cross(rc, u);
    j_wa = cross(ra, u);
    mass += mc + ma + ic * j_wc * j_wc + ia * j_wa * j_wa;
  }
  if (J.has(B2G_JF_REVOLUTE_B)) {
    j_vbd = v2<double>(0.0, 0.0);
    j_wb = 1.0;
    j_wd = 1.0;
    mass += sq(ratio) * (ib + id_);
  } else {
    const JV2 u = rmul(qd, J.g2(B2G_JD_PERP_X));
    const JV2 rd = rmul(qd, J.g2(B2G_JD_G1_X) - B.lc[3]);
    const JV2 rb = rmul(qb, J.g2(B2G_JD_LB_X) - B.lc[1]);
    j_vbd = u;
    j_wd = cross(rd, u);
    j_wb = cross(rb, u);
    mass += sq(ratio) * (md + mb) + id_ * sq(j_wd) + ib * j_wb * j_wb;
  }
  mass = mass > 0.0 ? 1.0 / mass : 0.0;
  J.s2(B2G_JD_GEA_JVAC_X, j_vac);
  J.s2(B2G_JD_GEA_JVBD_X, j_vbd);
  J.d[B2G_JD_GEA_JWA] = j_wa;
  J.d[B2G_JD_GEA_JWB] = j_wb;
  J.d[B2G_JD_GEA_JWC] = j_wc;
  J.d[B2G_JD_GEA_JWD] = j_wd;
  J.d[B2G_JD_GEA_MASS] = mass;
  if (st.warm) {  // no dt_ratio scaling here (gear.py:319-328)
    const double impulse = J.d[B2G_JD_IMP_X];
    Vs[0] = Vs[0] + scale(ma * impulse, j_vac);
    Ws[0] += ia * impulse * j_wa;
    Vs[1] = Vs[1] + scale(mb * impulse, j_vbd);
    Ws[1] += ib * impulse * j_wb;
    Vs[2] = Vs[2] - scale(mc * impulse, j_vac);
    Ws[2] -= ic * impulse * j_wc;
    Vs[3] = Vs[3] - scale(md * impulse, j_vbd);
    Ws[3] -= id_ * impulse * j_wd;
  } else {
    J.d[B2G_JD_IMP_X] = 0.0;
  }
  for (int k = 0; k < 4; ++k) {
    J.s2(B2G_JD_GEA_V + 2 * k, Vs[k]);
    J.s2(B2G_JD_GEA_C + 2 * k, Cs[k]);
    J.d[B2G_JD_GEA_A + k] = As[k];
    io.set_vel(B.idx[k], Vs[k], Ws[k]);
  }
}

template <class IO>
__device__ void gear_solve_velocity(const JRec& J, const JBodies& B, const IO& io) {
  JV2 Vs[4];
  double Ws[4];
  for (int k = 0; k < 4; ++k) {
    JV2 v;
    Vs[k] = J.g2(B2G_JD_GEA_V + 2 * k);
    io.vel(B.idx[k], v, Ws[k]);
  }
  const JV2 j_vac = J.g2(B2G_JD_GEA_JVAC_X), j_vbd = J.g2(B2G_JD_GEA_JVBD_X);
  const double j_wa = J.d[B2G_JD_GEA_JWA], j_wb = J.d[B2G_JD_GEA_JWB];
  const double j_wc = J.d[B2G_JD_GEA_JWC], j_wd = J.d[B2G_JD_GEA_JWD];
  double Cdot = dot(j_vac, Vs[0] - Vs[2]) + dot(j_vbd, Vs[1] - Vs[3]);
  Cdot += (j_wa * Ws[0] - j_wc * Ws[2]) + (j_wb * Ws[1] - j_wd * Ws[3]);
  const double impulse = -J.d[B2G_JD_GEA_MASS] * Cdot;
  J.d[B2G_JD_IMP_X] += impulse;
  const JV2 lvj[4] = {j_vac, j_vbd, -j_vac, -j_vbd};
  const double avj[4] = {j_wa, j_wb, -j_wc, -j_wd};
  for (int k = 0; k < 4; ++k) {
    Vs[k] = Vs[k] + scale(B.im[k] * impulse, lvj[k]);
    Ws[k] += (B.iI[k] * impulse) * avj[k];
  }
  for (int k = 0; k < 4; ++k) {
    J.s2(B2G_JD_GEA_V + 2 * k, Vs[k]);
    io.set_vel(B.idx[k], Vs[k], Ws[k]);
  }
}

template <class IO>
__device__ bool gear_solve_position(const JRec& J, const JBodies& B, const IO& io) {
  const double ratio = J.d[B2G_JD_RATIO];
  JV2 ca = J.g2(B2G_JD_GEA_C), cb = J.g2(B2G_JD_GEA_C + 2);
  JV2 cc = J.g2(B2G_JD_GEA_C + 4), cd = J.g2(B2G_JD_GEA_C + 6);
  double aa = J.d[B2G_JD_GEA_A], ab = J.d[B2G_JD_GEA_A + 1];
  double ac = J.d[B2G_JD_GEA_A + 2], ad = J.d[B2G_JD_GEA_A + 3];
  const Rot<double> qa = jrot(aa), qb = jrot(ab), qc = jrot(ac), qd = jrot(ad);
  const double ia = B.iI[0], ib = B.iI[1], ic = B.iI[2], id_ = B.iI[3];
  const double ma = B.im[0], mb = B.im[1], mc = B.im[2], md = B.im[3];
  double mass = 0.0, coordinate_a, coordinate_b;
  JV2 j_vac, j_vbd;
  double j_wa, j_wb, j_wc, j_wd;
  if (J.has(B2G_JF_REVOLUTE_A)) {
    j_vac = v2<double>(0.0, 0.0);
    j_wa = 1.0;
    j_wc = 1.0;
    mass += ia + ic;
    coordinate_a = aa - ac - J.d[B2G_JD_REF];
  } else {
    const JV2 axis_c = J.g2(B2G_JD_AXIS_X);
    const JV2 u = rmul(qc, axis_c);
    const JV2 rc = rmul(qc, J.g2(B2G_JD_G0_X) - B.lc[2]);
    const JV2 ra = rmul(qa, J.g2(B2G_JD_LA_X) - B.lc[0]);
    j_vac = u;
    j_wc = cross(rc, u);
    j_wa = cross(ra, u);
    mass += mc + ma + ic * sq(j_wc) + ia * sq(j_wa);
    const JV2 pc = J.g2(B2G_JD_G0_X) - B.lc[2];
    const JV2 pa = rmul_t(qc, ra + (ca - cc));
    coordinate_a = dot(pa - pc, axis_c);
  }
  if (J.has(B2G_JF_REVOLUTE_B)) {
    j_vbd = v2<double>(0.0, 0.0);
    j_wb = 1.0;
    j_wd = 1.0;
    mass += ib + id_;
    coordinate_b = ab - ad - J.d[B2G_JD_REF_B];
  } else {
    const JV2 axis_d = J.g2(B2G_JD_PERP_X);
    const JV2 u = rmul(qd, axis_d);
    const JV2 rd = rmul(qd, J.g2(B2G_JD_G1_X) - B.lc[3]);
    const JV2 rb = rmul(qb, J.g2(B2G_JD_LB_X) - B.lc[1]);
    j_vbd = scale(ratio, u);
    j_wd = ratio * cross(rd, u);
    j_wb = ratio * cross(rb, u);
    mass += sq(ratio) * (md + mb) + id_ * sq(j_wd) + ib * sq(j_wb);
    const JV2 pd = J.g2(B2G_JD_G1_X) - B.lc[3];
    const JV2 pb = rmul_t(qd, rb + (cb - cd));
    coordinate_b = dot(pb - pd, axis_d);
  }
  const double C = (coordinate_a + ratio * coordinate_b) - J.d[B2G_JD_LENGTH];
  if (mass > 0.0) {
    const double impulse = -C / mass;
    ca = ca + scale(ma * impulse, j_vac);
    cb = cb + scale(mb * impulse, j_vbd);
    cc = cc - scale(mc * impulse, j_vac);
    cd = cd - scale(md * impulse, j_vbd);
    aa += ia * impulse * j_wa;
    ab += ib * impulse * j_wb;
    ac -= ic * impulse * j_wc;
    ad -= id_ * impulse * j_wd;
    io.set_pos(B.idx[0], ca, aa);
    io.set_pos(B.idx[1], cb, ab);
    io.set_pos(B.idx[2], cc, ac);
    io.set_pos(B.idx[3], cd, ad);
  }
  return true;
}

// ---- wheel (wheel.py:241-417) ----
template <class IO>
__device__ void wheel_init(const JRec& J, const JBodies& B, const IO& io, const JStep& st) {
  const double ma = B.im[0], mb = B.im[1], ia = B.iI[0], ib = B.iI[1];
  JState s;
  jload_pos(io, B, s);
  jload_vel(io, B, s);
  const Rot<double> qa = jrot(s.aa);
  const JV2 ra = rmul(qa, J.g2(B2G_JD_LA_X) - B.lc[0]);
  const JV2 rb = lever(s.ab, J.g2(B2G_JD_LB_X), B.lc[1]);
  J.s2(B2G_JD_RA_X, ra);
  J.s2(B2G_JD_RB_X, rb);
  const JV2 d = s.cb + rb - s.ca - ra;
  const JV2 ay = rmul(qa, J.g2(B2G_JD_PERP_X));
  J.s2(B2G_JD_WHE_AY_X, ay);
  const double s_ay = cross(d + ra, ay), s_by = cross(rb, ay);
  J.d[B2G_JD_WHE_S_AY] = s_ay;
  J.d[B2G_JD_WHE_S_BY] = s_by;
  double mass = ma + mb + ia * sq(s_ay) + ib * sq(s_by);
  if (mass > 0.0) mass = 1.0 / mass;
  J.d[B2G_JD_WHE_MASS] = mass;
  J.d[B2G_JD_WHE_SPRING_MASS] = 0.0;
  J.d[B2G_JD_WHE_BIAS] = 0.0;
  J.d[B2G_JD_WHE_GAMMA] = 0.0;
  if (J.d[B2G_JD_FREQ] > 0.0) {
    const JV2 ax = rmul(qa, J.g2(B2G_JD_AXIS_X));
    J.s2(B2G_JD_WHE_AX_X, ax);
    const double s_ax = cross(d + ra, ax), s_bx = cross(rb, ax);
    J.d[B2G_JD_WHE_S_AX] = s_ax;
    J.d[B2G_JD_WHE_S_BX] = s_bx;
    const double inv_mass = ma + mb + ia * sq(s_ax) + ib * sq(s_bx);
    if (inv_mass > 0.0) {
      double spring_mass = 1.0 / inv_mass;
      const double C = dot(d, ax);
      const double omega = 2.0 * kJPi * J.d[B2G_JD_FREQ];
      const double dd = 2.0 * spring_mass * J.d[B2G_JD_DAMP] * omega;
      const double k = spring_mass * sq(omega);
      const double dt = st.dt;
      double gamma = dt * (dd + dt * k);
      if (gamma > 0.0) gamma = 1.0 / gamma;
      J.d[B2G_JD_WHE_GAMMA] = gamma;
      J.d[B2G_JD_WHE_BIAS] = C * dt * k * gamma;
      spring_mass = inv_mass + gamma;
      if (spring_mass > 0.0) spring_mass = 1.0 / spring_mass;
      J.d[B2G_JD_WHE_SPRING_MASS] = spring_mass;
    }
  } else {
    J.d[B2G_JD_SPRING_IMP] = 0.0;
  }
  if (J.has(B2G_JF_MOTOR)) {
    double motor_mass = ia + ib;
    if (motor_mass > 0.0) motor_mass = 1.0 / motor_mass;
    J.d[B2G_JD_WHE_MOTOR_MASS] = motor_mass;
  } else {
    J.d[B2G_JD_WHE_MOTOR_MASS] = 0.0;
    J.d[B2G_JD_MOTOR_IMP] = 0.0;
  }
  if (st.warm) {
    const double r = st.dt_ratio;
    const double impulse = J.d[B2G_JD_IMP_X] * r;
    const double spring_impulse = J.d[B2G_JD_SPRING_IMP] * r;
    const double motor_impulse = J.d[B2G_JD_MOTOR_IMP] * r;
    J.d[B2G_JD_IMP_X] = impulse;
    J.d[B2G_JD_SPRING_IMP] = spring_impulse;
    J.d[B2G_JD_MOTOR_IMP] = motor_impulse;
    // the spring terms use the stored axis and arms (stale when the frequency is 0)
    const JV2 P = scale(impulse, ay) + scale(spring_impulse, J.g2(B2G_JD_WHE_AX_X));
    const double La = impulse * s_ay + spring_impulse * J.d[B2G_JD_WHE_S_AX] + motor_impulse;
    const double Lb = impulse * s_by + spring_impulse * J.d[B2G_JD_WHE_S_BX] + motor_impulse;
    s.va = s.va - scale(ma, P);
    s.wa -= ia * La;
    s.vb = s.vb + scale(mb, P);
    s.wb += ib * Lb;
  } else {
    J.d[B2G_JD_IMP_X] = 0.0;
    J.d[B2G_JD_SPRING_IMP] = 0.0;
    J.d[B2G_JD_MOTOR_IMP] = 0.0;
  }
  jstore_vel(io, B, s);
}

template <class IO>
__device__ void wheel_solve_velocity(const JRec& J, const JBodies& B, const IO& io, const JStep& st) {
  const double ma = B.im[0], mb = B.im[1], ia = B.iI[0], ib = B.iI[1];
  JState s;
  jload_vel(io, B, s);
  const JV2 ax = J.g2(B2G_JD_WHE_AX_X), ay = J.g2(B2G_JD_WHE_AY_X);
  const double s_ax = J.d[B2G_JD_WHE_S_AX], s_bx = J.d[B2G_JD_WHE_S_BX];
  const double s_ay = J.d[B2G_JD_WHE_S_AY], s_by = J.d[B2G_JD_WHE_S_BY];
  double Cdot = dot(ax, s.vb - s.va) + s_bx * s.wb - s_ax * s.wa;
  double impulse = -J.d[B2G_JD_WHE_SPRING_MASS] *
                   (Cdot + J.d[B2G_JD_WHE_BIAS] + J.d[B2G_JD_WHE_GAMMA] * J.d[B2G_JD_SPRING_IMP]);
  J.d[B2G_JD_SPRING_IMP] += impulse;
  JV2 P = scale(impulse, ax);
  double La = impulse * s_ax, Lb = impulse * s_bx;
  s.va = s.va - scale(ma, P);
  s.wa -= ia * La;
  s.vb = s.vb + scale(mb, P);
  s.wb += ib * Lb;
  Cdot = s.wb - s.wa - J.d[B2G_JD_MOTOR_SPEED];
  impulse = -J.d[B2G_JD_WHE_MOTOR_MASS] * Cdot;
  const double old = J.d[B2G_JD_MOTOR_IMP];
  const double max_impulse = st.dt * J.d[B2G_JD_MAX_MOTOR];
  J.d[B2G_JD_MOTOR_IMP] = pclamp(old + impulse, -max_impulse, max_impulse);
  impulse = J.d[B2G_JD_MOTOR_IMP] - old;
  s.wa -= ia * impulse;
  s.wb += ib * impulse;
  Cdot = dot(ay, s.vb - s.va) + s_by * s.wb - s_ay * s.wa;
  impulse = -J.d[B2G_JD_WHE_MASS] * Cdot;
  J.d[B2G_JD_IMP_X] += impulse;
  P = scale(impulse, ay);
  La = impulse * s_ay;
  Lb = impulse * s_by;
  s.va = s.va - scale(ma, P);
  s.wa -= ia * La;
  s.vb = s.vb + scale(mb, P);
  s.wb += ib * Lb;
  jstore_vel(io, B, s);
}

template <class IO>
__device__ bool wheel_solve_position(const JRec& J, const JBodies& B, const IO& io) {
  const double ma = B.im[0], mb = B.im[1], ia = B.iI[0], ib = B.iI[1];
  JState s;
  jload_pos(io, B, s);
  const Rot<double> qa = jrot(s.aa);
  const JV2 ra = rmul(qa, J.g2(B2G_JD_LA_X) - B.lc[0]);
  const JV2 rb = lever(s.ab, J.g2(B2G_JD_LB_X), B.lc[1]);
  const JV2 d = (s.cb - s.ca) + rb - ra;
  // the effective mass from the setup's arms s_ay, s_by (wheel.py:402)
  const double k = ma + mb + ia * sq(J.d[B2G_JD_WHE_S_AY]) + ib * sq(J.d[B2G_JD_WHE_S_BY]);
  const JV2 ay = rmul(qa, J.g2(B2G_JD_PERP_X));
  const double s_ay = cross(d + ra, ay), s_by = cross(rb, ay);
  const double C = dot(d, ay);
  const double impulse = k != 0.0 ? -C / k : 0.0;
  const JV2 P = scale(impulse, ay);
  const double La = impulse * s_ay, Lb = impulse * s_by;
  s.ca = s.ca - scale(ma, P);
  s.aa -= ia * La;
  s.cb = s.cb + scale(mb, P);
  s.ab += ib * Lb;
  jstore_pos(io, B, s);
  return fabs(C) <= kJLinearSlop;
}

// ---- weld (weld.py:127-257): the setup's mass matrix has col2.z = c3x (weld.py:162); the
// position solve rebuilds it with c3y and stores it in the joint ----
template <class IO>
__device__ void weld_init(const JRec& J, const JBodies& B, const IO& io, const JStep& st) {
  const double ma = B.im[0], mb = B.im[1], ia = B.iI[0], ib = B.iI[1];
  JState s;
  jload_pos(io, B, s);
  jload_vel(io, B, s);
  const JV2 ra = lever(s.aa, J.g2(B2G_JD_LA_X), B.lc[0]);
  const JV2 rb = lever(s.ab, J.g2(B2G_JD_LB_X), B.lc[1]);
  J.s2(B2G_JD_RA_X, ra);
  J.s2(B2G_JD_RB_X, rb);
  const double c2x = -ra.y * ra.x * ia - rb.y * rb.x * ib;
  const double c3x = -ra.y * ia - rb.y * ib;
  JM33 m;
  m.c1 = jv3(ma + mb + ra.y * ra.y * ia + rb.y * rb.y * ib, c2x, c3x);
  m.c2 = jv3(c2x, ma + mb + ra.x * ra.x * ia + rb.x * rb.x * ib, c3x);
  m.c3 = jv3(c3x, ra.x * ia + rb.x * ib, ia + ib);
  J.s33(B2G_JD_WEL_MASS, m);
  if (st.warm) {
    const JV3 imp = scale3(st.dt_ratio, J.g3(B2G_JD_IMP_X));
    J.s3(B2G_JD_IMP_X, imp);
    const JV2 P = v2<double>(imp.x, imp.y);
    s.va = s.va - scale(ma, P);
    s.wa -= ia * (cross(ra, P) + imp.z);
    s.vb = s.vb + scale(mb, P);
    s.wb += ib * (cross(rb, P) + imp.z);
  } else {
    J.s3(B2G_JD_IMP_X, jv3(0.0, 0.0, 0.0));
  }
  jstore_vel(io, B, s);
}

template <class IO>
__device__ void weld_solve_velocity(const JRec& J, const JBodies& B, const IO& io) {
  JState s;
  jload_vel(io, B, s);
  const JV2 ra = J.g2(B2G_JD_RA_X), rb = J.g2(B2G_JD_RB_X);
  const JV2 Cdot1 = s.vb + cross_sv(s.wb, rb) - s.va - cross_sv(s.wa, ra);
  const double Cdot2 = s.wb - s.wa;
  const JV3 impulse = solve33(J.g33(B2G_JD_WEL_MASS), neg3(jv3(Cdot1.x, Cdot1.y, Cdot2)));
  J.s3(B2G_JD_IMP_X, add3(J.g3(B2G_JD_IMP_X), impulse));
  const JV2 P = v2<double>(impulse.x, impulse.y);
  s.va = s.va - scale(B.im[0], P);
  s.wa -= B.iI[0] * (cross(ra, P) + impulse.z);
  s.vb = s.vb + scale(B.im[1], P);
  s.wb += B.iI[1] * (cross(rb, P) + impulse.z);
  jstore_vel(io, B, s);
}

template <class IO>
__device__ bool weld_solve_position(const JRec& J, const JBodies& B, const IO& io) {
  const double ma = B.im[0], mb = B.im[1], ia = B.iI[0], ib = B.iI[1];
  JState s;
  jload_pos(io, B, s);
  const JV2 ra = lever(s.aa, J.g2(B2G_JD_LA_X), B.lc[0]);
  const JV2 rb = lever(s.ab, J.g2(B2G_JD_LB_X), B.lc[1]);
  const JV2 C1 = s.cb + rb - s.ca - ra;
  const double C2 = s.ab - s.aa - J.d[B2G_JD_REF];
  const double position_error = jlen(C1), angular_error = fabs(C2);
  const double c2x = -ra.y * ra.x * ia - rb.y * rb.x * ib;
  const double c3x = -ra.y * ia - rb.y * ib;
  const double c3y = ra.x * ia + rb.x * ib;
  JM33 m;
  m.c1 = jv3(ma + mb + ra.y * ra.y * ia + rb.y * rb.y * ib, c2x, c3x);
  m.c2 = jv3(c2x, ma + mb + ra.x * ra.x * ia + rb.x * rb.x * ib, c3y);
  m.c3 = jv3(-ra.y * ia - rb.y * ib, ra.x * ia + rb.x * ib, ia + ib);
  J.s33(B2G_JD_WEL_MASS, m);
  const JV3 impulse = neg3(solve33(m, jv3(C1.x, C1.y, C2)));
  const JV2 P = v2<double>(impulse.x, impulse.y);
  s.ca = s.ca - scale(ma, P);
  s.aa -= ia * (cross(ra, P) + impulse.z);
  s.cb = s.cb + scale(mb, P);
  s.ab += ib * (cross(rb, P) + impulse.z);
  jstore_pos(io, B, s);
  return position_error <= kJLinearSlop && angular_error <= kJAngularSlop;
}

// ---- friction (friction.py:115-233): the velocity solve stores its results into the position
// list (friction.py:228-229), so only its impulses act, through the next step's warm start ----
template <class IO>
__device__ void friction_init(const JRec& J, const JBodies& B, const IO& io, const JStep& st) {
  const double ma = B.im[0], mb = B.im[1], ia = B.iI[0], ib = B.iI[1];
  JState s;
  jload_pos(io, B, s);
  jload_vel(io, B, s);
  const JV2 ra = lever(s.aa, J.g2(B2G_JD_LA_X), B.lc[0]);
  const JV2 rb = lever(s.ab, J.g2(B2G_JD_LB_X), B.lc[1]);
  J.s2(B2G_JD_RA_X, ra);
  J.s2(B2G_JD_RB_X, rb);
  M22<double> K;
  K.c1x = ma + mb + ia * ra.y * ra.y + ib * rb.y * rb.y;
  K.c1y = -ia * ra.x * ra.y - ib * rb.x * rb.y;
  K.c2x = K.c1y;
  K.c2y = ma + mb + ia * ra.x * ra.x + ib * rb.x * rb.x;
  J.s22(B2G_JD_FRI_LMASS, m22_inverse(K));
  double angular_mass = ia + ib;
  if (angular_mass > 0.0) angular_mass = 1.0 / angular_mass;
  J.d[B2G_JD_FRI_AMASS] = angular_mass;
  if (st.warm) {
    const JV2 li = scale(st.dt_ratio, J.g2(B2G_JD_IMP_X));
    const double ang = J.d[B2G_JD_IMP_Z] * st.dt_ratio;
    J.s2(B2G_JD_IMP_X, li);
    J.d[B2G_JD_IMP_Z] = ang;
    s.va = s.va - scale(ma, li);
    s.wa -= ia * (cross(ra, li) + ang);
    s.vb = s.vb + scale(mb, li);
    s.wb += ib * (cross(rb, li) + ang);
  } else {
    J.s2(B2G_JD_IMP_X, v2<double>(0.0, 0.0));
    J.d[B2G_JD_IMP_Z] = 0.0;
  }
  jstore_vel(io, B, s);
}

template <class IO>
__device__ void friction_solve_velocity(const JRec& J, const JBodies& B, const IO& io, const JStep& st) {
  JState s;
  jload_vel(io, B, s);
  const JV2 ra = J.g2(B2G_JD_RA_X), rb = J.g2(B2G_JD_RB_X);
  const double dt = st.dt;
  const double Cdot = s.wb - s.wa;
  double impulse = -J.d[B2G_JD_FRI_AMASS] * Cdot;
  const double old = J.d[B2G_JD_IMP_Z];
  const double max_impulse = dt * J.d[B2G_JD_MAX_TORQUE];
  J.d[B2G_JD_IMP_Z] = pclamp(old + impulse, -max_impulse, max_impulse);
  impulse = J.d[B2G_JD_IMP_Z] - old;
  s.wa -= B.iI[0] * impulse;
  s.wb += B.iI[1] * impulse;
  const JV2 Cdot2 = s.vb + cross_sv(s.wb, rb) - s.va - cross_sv(s.wa, ra);
  const JV2 limp = -mmul(J.g22(B2G_JD_FRI_LMASS), Cdot2);
  JV2 li = J.g2(B2G_JD_IMP_X) + limp;
  const double lmax = dt * J.d[B2G_JD_MAX_MOTOR];
  if (jlen2(li) > sq(lmax)) li = scale(lmax, normalized(li));
  J.s2(B2G_JD_IMP_X, li);
  // the updated velocities are dropped (friction.py:228-229)
}

// ---- rope (rope.py:104-248): the attribute _u keeps the unnormalised separation (`u /= length`
// rebinds the local u), and the warm start and the velocity solve use it ----
template <class IO>
__device__ void rope_init(const JRec& J, const JBodies& B, const IO& io, const JStep& st) {
  const double ma = B.im[0], mb = B.im[1], ia = B.iI[0], ib = B.iI[1];
  JState s;
  jload_pos(io, B, s);
  jload_vel(io, B, s);
  const JV2 ra = lever(s.aa, J.g2(B2G_JD_LA_X), B.lc[0]);
  const JV2 rb = lever(s.ab, J.g2(B2G_JD_LB_X), B.lc[1]);
  J.s2(B2G_JD_RA_X, ra);
  J.s2(B2G_JD_RB_X, rb);
  JV2 u = s.cb + rb - s.ca - ra;
  J.s2(B2G_JD_ROP_U_X, u);
  const double length = jlen(u);
  J.d[B2G_JD_ROP_LENGTH] = length;
  const double C = length - J.d[B2G_JD_LENGTH];
  J.i[B2G_JI_STATE] = C > 0.0 ? B2G_LIMIT_AT_UPPER : B2G_LIMIT_INACTIVE;
  if (length > kJLinearSlop) {
    u = vdiv(u, length);
  } else {
    J.d[B2G_JD_ROP_MASS] = 0.0;
    J.d[B2G_JD_IMP_X] = 0.0;
    return;
  }
  const double cra = cross(ra, u), crb = cross(rb, u);
  const double inv_mass = ma + ia * cra * cra + mb + ib * crb * crb;
  J.d[B2G_JD_ROP_MASS] = inv_mass != 0.0 ? 1.0 / inv_mass : 0.0;
  if (st.warm) {
    const double imp = J.d[B2G_JD_IMP_X] * st.dt_ratio;
    J.d[B2G_JD_IMP_X] = imp;
    const JV2 P = scale(imp, J.g2(B2G_JD_ROP_U_X));
    s.va = s.va - scale(ma, P);
    s.wa -= ia * cross(ra, P);
    s.vb = s.vb + scale(mb, P);
    s.wb += ib * cross(rb, P);
  } else {
    J.d[B2G_JD_IMP_X] = 0.0;
  }
  jstore_vel(io, B, s);
}

template <class IO>
__device__ void rope_solve_velocity(const JRec& J, const JBodies& B, const IO& io, const JStep& st) {
  JState s;
  jload_vel(io, B, s);
  const JV2 ra = J.g2(B2G_JD_RA_X), rb = J.g2(B2G_JD_RB_X), u = J.g2(B2G_JD_ROP_U_X);
  const JV2 vpa = s.va + cross_sv(s.wa, ra), vpb = s.vb + cross_sv(s.wb, rb);
  const double C = J.d[B2G_JD_ROP_LENGTH] - J.d[B2G_JD_LENGTH];
  double Cdot = dot(u, vpb - vpa);
  if (C < 0.0) Cdot += st.inv_dt * C;
  double impulse = -J.d[B2G_JD_ROP_MASS] * Cdot;
  const double old = J.d[B2G_JD_IMP_X];
  J.d[B2G_JD_IMP_X] = pmin(0.0, old + impulse);
  impulse = J.d[B2G_JD_IMP_X] - old;
  const JV2 P = scale(impulse, u);
  s.va = s.va - scale(B.im[0], P);
  s.wa -= B.iI[0] * cross(ra, P);
  s.vb = s.vb + scale(B.im[1], P);
  s.wb += B.iI[1] * cross(rb, P);
  jstore_vel(io, B, s);
}

template <class IO>
__device__ bool rope_solve_position(const JRec& J, const JBodies& B, const IO& io) {
  JState s;
  jload_pos(io, B, s);
  const JV2 ra = lever(s.aa, J.g2(B2G_JD_LA_X), B.lc[0]);
  const JV2 rb = lever(s.ab, J.g2(B2G_JD_LB_X), B.lc[1]);
  JV2 u = s.cb + rb - s.ca - ra;
  const double length = jlen(u);
  u = normalized(u);
  double C = length - J.d[B2G_JD_LENGTH];
  C = pclamp(C, 0.0, kJMaxLinearCorrection);
  const double impulse = -J.d[B2G_JD_ROP_MASS] * C;
  const JV2 P = scale(impulse, u);
  s.ca = s.ca - scale(B.im[0], P);
  s.aa -= B.iI[0] * cross(ra, P);
  s.cb = s.cb + scale(B.im[1], P);
  s.ab += B.iI[1] * cross(rb, P);
  jstore_pos(io, B, s);
  return length - J.d[B2G_JD_LENGTH] < kJLinearSlop;
}

// Joint._init_velocity_constraints; returns false when a mouse joint's spring is too weak (the
// reference asserts, mouse.py:173).
template <class IO>
__device__ bool joint_init(const JRec& J, const JBodies& B, const IO& io, const JStep& st) {
  switch (J.i[B2G_JI_TYPE]) {
    case B2G_JT_REVOLUTE: revolute_init(J, B, io, st); break;
    case B2G_JT_PRISMATIC: prismatic_init(J, B, io, st); break;
    case B2G_JT_DISTANCE: distance_init(J, B, io, st); break;
    case B2G_JT_PULLEY: pulley_init(J, B, io, st); break;
    case B2G_JT_MOUSE: return mouse_init(J, B, io, st);
    case B2G_JT_GEAR: gear_init(J, B, io, st); break;
    case B2G_JT_WHEEL: wheel_init(J, B, io, st); break;
    case B2G_JT_WELD: weld_init(J, B, io, st); break;
    case B2G_JT_FRICTION: friction_init(J, B, io, st); break;
    case B2G_JT_ROPE: rope_init(J, B, io, st); break;
    default: break;
  }
  return true;
}
template <class IO>
__device__ void joint_solve_velocity(const JRec& J, const JBodies& B, const IO& io, const JStep& st) {
  switch (J.i[B2G_JI_TYPE]) {
    case B2G_JT_REVOLUTE: revolute_solve_velocity(J, B, io, st); break;
    case B2G_JT_PRISMATIC: prismatic_solve_velocity(J, B, io, st); break;
    case B2G_JT_DISTANCE: distance_solve_velocity(J, B, io); break;
    case B2G_JT_PULLEY: pulley_solve_velocity(J, B, io); break;
    case B2G_JT_MOUSE: mouse_solve_velocity(J, B, io, st); break;
    case B2G_JT_GEAR: gear_solve_velocity(J, B, io); break;
    case B2G_JT_WHEEL: wheel_solve_velocity(J, B, io, st); break;
    case B2G_JT_WELD: weld_solve_velocity(J, B, io); break;
    case B2G_JT_FRICTION: friction_solve_velocity(J, B, io, st); break;
    case B2G_JT_ROPE: rope_solve_velocity(J, B, io, st); break;
    default: break;
  }
}
template <class IO>
__device__ bool joint_solve_position(const JRec& J, const JBodies& B, const IO& io) {
  switch (J.i[B2G_JI_TYPE]) {
    case B2G_JT_REVOLUTE: return revolute_solve_position(J, B, io);
    case B2G_JT_PRISMATIC: return prismatic_solve_position(J, B, io);
    case B2G_JT_DISTANCE: return distance_solve_position(J, B, io);
    case B2G_JT_PULLEY: return pulley_solve_position(J, B, io);
    case B2G_JT_GEAR: return gear_solve_position(J, B, io);
    case B2G_JT_WHEEL: return wheel_solve_position(J, B, io);
    case B2G_JT_WELD: return weld_solve_position(J, B, io);
    case B2G_JT_ROPE: return rope_solve_position(J, B, io);
    default: return true;  // mouse, friction (mouse.py:228-229, friction.py:231-233)
  }
}

}  // namespace b2g
