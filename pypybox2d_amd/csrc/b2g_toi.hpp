// b2g_toi.hpp -- continuous-physics pieces, FP64: the GJK distance of two shapes, the time of
// impact of two sweeps, and the sweep operations of a body.
//
// They follow pypybox2d operation by operation:
//   DistanceProxy / Simplex / shape_distance   distance.py:37-561 (radii not applied: the
//                                              time of impact asks with use_radii False)
//   SeparationFunction / time_of_impact        toi.py:44-391
//   Sweep.get_transform / advance / normalize  common.py:895-923
// Reference behaviour kept as it is: a one-point separation axis stays unnormalised
// (toi.py:68-71 discards the unit vector), and ties in a support search take the first
// vertex.  One bound is added: the reference's inner loop of time_of_impact has no limit on
// its push-back iterations (toi.py:335-339 leaves it commented out); here it stops after
// kToiPushBackLimit and reports the pair as stuck (the caller flags the world).
#pragma once

#include "b2g_collide.hpp"

namespace b2g {

constexpr int kDistanceMaxIters = 20;    // settings.py:129
constexpr int kToiMaxIters = 20;         // settings.py:82
constexpr int kToiPushBackLimit = 1000;
constexpr double kToiLinearSlop = 0.005;  // settings.py:61
constexpr double kToiPi = 3.141592653589793;

enum : int { TOI_UNKNOWN = 0, TOI_FAILED, TOI_OVERLAPPED, TOI_TOUCHING, TOI_SEPARATED, TOI_STUCK };

// DistanceProxy (distance.py:355-400): a circle's centre, a polygon's vertices, an edge's ends.
struct TProxy {
  V2<double> v[8];
  int n;
  double radius;
};

__device__ inline void tproxy_init(TProxy& p, const Geo<double, double>& g) {
  p.radius = g.radius();
  if (g.type == SHAPE_CIRCLE) {
    p.v[0] = g.cp();
    p.n = 1;
  } else if (g.type == SHAPE_POLYGON) {
    p.n = g.n;
    for (int i = 0; i < g.n; ++i) p.v[i] = g.v(i);
  } else {
    p.v[0] = g.e(1);
    p.v[1] = g.e(2);
    p.n = 2;
  }
}

// get_support: the first vertex of largest dot (distance.py:402-405)
__device__ inline int tproxy_support(const TProxy& p, V2<double> d) {
  int best = 0;
  double bv = dot(p.v[0], d);
  for (int i = 1; i < p.n; ++i) {
    const double x = dot(p.v[i], d);
    if (x > bv) {
      bv = x;
      best = i;
    }
  }
  return best;
}

struct SimplexVertex {
  V2<double> wa, wb, w;
  double a;
  int ia, ib;
};
struct GjkSimplex {
  SimplexVertex v[3];
  int count;
};
struct GjkCache {
  double metric;
  int count;
  int ia[3], ib[3];
};

__device__ inline double tlength(V2<double> a) { return sqrt(a.x * a.x + a.y * a.y); }

__device__ inline double gjk_metric(const GjkSimplex& s) {  // distance.py:180-193
  if (s.count == 1) return 0.0;
  if (s.count == 2) return tlength(s.v[0].w - s.v[1].w);
  return cross(s.v[1].w - s.v[0].w, s.v[2].w - s.v[0].w);
}

__device__ inline void gjk_vertex(SimplexVertex& v, const TProxy& pa, const XF<double>& xfa, const TProxy& pb,
                                  const XF<double>& xfb) {
  v.wa = xmul(xfa, pa.v[v.ia]);
  v.wb = xmul(xfb, pb.v[v.ib]);
  v.w = v.wb - v.wa;
}

__device__ inline void gjk_solve2(GjkSimplex& s) {  // distance.py:195-245
  const V2<double> w1 = s.v[0].w, w2 = s.v[1].w, e12 = w2 - w1;
  const double d12_2 = -dot(w1, e12);
  if (d12_2 <= 0.0) {
    s.v[0].a = 1.0;
    s.count = 1;
    return;
  }
  const double d12_1 = dot(w2, e12);
  if (d12_1 <= 0.0) {
    s.v[1].a = 1.0;
    s.count = 1;
    s.v[0] = s.v[1];
    return;
  }
  const double inv = 1.0 / (d12_1 + d12_2);
  s.v[0].a = d12_1 * inv;
  s.v[1].a = d12_2 * inv;
  s.count = 2;
}

__device__ inline void gjk_solve3(GjkSimplex& s) {  // distance.py:247-353
  const V2<double> w1 = s.v[0].w, w2 = s.v[1].w, w3 = s.v[2].w;
  const V2<double> e12 = w2 - w1, e13 = w3 - w1, e23 = w3 - w2;
  const double d12_1 = dot(w2, e12), d12_2 = -dot(w1, e12);
  const double d13_1 = dot(w3, e13), d13_2 = -dot(w1, e13);
  const double d23_1 = dot(w3, e23), d23_2 = -dot(w2, e23);
  const double n123 = cross(e12, e13);
  const double d123_1 = n123 * cross(w2, w3), d123_2 = n123 * cross(w3, w1), d123_3 = n123 * cross(w1, w2);
  if (d12_2 <= 0.0 && d13_2 <= 0.0) {
    s.v[0].a = 1.0;
    s.count = 1;
    return;
  }
  if (d12_1 > 0.0 && d12_2 > 0.0 && d123_3 <= 0.0) {
    const double inv = 1.0 / (d12_1 + d12_2);
    s.v[0].a = d12_1 * inv;
    s.v[1].a = d12_2 * inv;
    s.count = 2;
    return;
  }
  if (d13_1 > 0.0 && d13_2 > 0.0 && d123_2 <= 0.0) {
    const double inv = 1.0 / (d13_1 + d13_2);
    s.v[0].a = d13_1 * inv;
    s.v[2].a = d13_2 * inv;
    s.count = 2;
    s.v[1] = s.v[2];
    return;
  }
  if (d12_1 <= 0.0 && d23_2 <= 0.0) {
    s.v[1].a = 1.0;
    s.count = 1;
    s.v[0] = s.v[1];
    return;
  }
  if (d13_1 <= 0.0 && d23_1 <= 0.0) {
    s.v[2].a = 1.0;
    s.count = 1;
    s.v[0] = s.v[2];
    return;
  }
  if (d23_1 > 0.0 && d23_2 > 0.0 && d123_1 <= 0.0) {
    const double inv = 1.0 / (d23_1 + d23_2);
    s.v[1].a = d23_1 * inv;
    s.v[2].a = d23_2 * inv;
    s.count = 2;
    s.v[0] = s.v[2];
    return;
  }
  const double inv = 1.0 / (d123_1 + d123_2 + d123_3);
  s.v[0].a = d123_1 * inv;
  s.v[1].a = d123_2 * inv;
  s.v[2].a = d123_3 * inv;
  s.count = 3;
}

// shape_distance with use_radii False (distance.py:424-561); the cache is read and written.
__device__ inline double gjk_distance(const TProxy& pa, const TProxy& pb, const XF<double>& xfa,
                                      const XF<double>& xfb, GjkCache& cache) {
  GjkSimplex s;
  // read_cache (distance.py:88-120)
  s.count = cache.count;
  for (int i = 0; i < s.count; ++i) {
    s.v[i].ia = cache.ia[i];
    s.v[i].ib = cache.ib[i];
    gjk_vertex(s.v[i], pa, xfa, pb, xfb);
    s.v[i].a = 0.0;
  }
  if (s.count > 1) {
    const double m1 = cache.metric, m2 = gjk_metric(s);
    if (m2 < 0.5 * m1 || 2.0 * m1 < m2 || m2 < NormEps<double>::value) s.count = 0;
  }
  if (s.count == 0) {
    s.v[0].ia = 0;
    s.v[0].ib = 0;
    gjk_vertex(s.v[0], pa, xfa, pb, xfb);
    s.v[0].a = 0.0;
    s.count = 1;
  }
  int iter = 0;
  while (iter < kDistanceMaxIters) {
    const int save_n = s.count;
    int save_a[3] = {0, 0, 0}, save_b[3] = {0, 0, 0};
    for (int i = 0; i < save_n; ++i) {
      save_a[i] = s.v[i].ia;
      save_b[i] = s.v[i].ib;
    }
    if (s.count == 2) {
      gjk_solve2(s);
    } else if (s.count == 3) {
      gjk_solve3(s);
      break;
    }
    // search direction (distance.py:140-154)
    V2<double> d;
    if (s.count == 1) {
      d = -s.v[0].w;
    } else {
      const V2<double> e12 = s.v[1].w - s.v[0].w;
      d = cross(e12, -s.v[0].w) > 0.0 ? cross_sv(1.0, e12) : cross_vs(e12, 1.0);
    }
    if (d.x * d.x + d.y * d.y < NormEps<double>::value * NormEps<double>::value) break;
    SimplexVertex& v = s.v[s.count];
    v.ia = tproxy_support(pa, rmul_t(xfa.q, -d));
    v.wa = xmul(xfa, pa.v[v.ia]);
    v.ib = tproxy_support(pb, rmul_t(xfb.q, d));
    v.wb = xmul(xfb, pb.v[v.ib]);
    v.w = v.wb - v.wa;
    ++iter;
    bool dup = false;
    for (int i = 0; i < save_n; ++i) dup |= save_a[i] == v.ia && save_b[i] == v.ib;
    if (dup) break;
    ++s.count;
  }
  // witness points (distance.py:156-178)
  V2<double> wa, wb;
  if (s.count == 1) {
    wa = s.v[0].wa;
    wb = s.v[0].wb;
  } else if (s.count == 2) {
    wa = scale(s.v[0].a, s.v[0].wa) + scale(s.v[1].a, s.v[1].wa);
    wb = scale(s.v[0].a, s.v[0].wb) + scale(s.v[1].a, s.v[1].wb);
  } else {
    wa = scale(s.v[0].a, s.v[0].wa) + scale(s.v[1].a, s.v[1].wa) + scale(s.v[2].a, s.v[2].wa);
    wb = wa;
  }
  const double dist = tlength(wa - wb);
  cache.metric = gjk_metric(s);
  cache.count = s.count;
  for (int i = 0; i < s.count; ++i) {
    cache.ia[i] = s.v[i].ia;
    cache.ib[i] = s.v[i].ib;
  }
  return dist;
}

// A body's sweep (common.py:870-923): centre and angle at the sweep's start (c0, a0, at
// fraction alpha0 of the step) and end (c, a), local centre lc.
struct TSweep {
  V2<double> c0, c, lc;
  double a0, a, alpha0;
};

__device__ inline XF<double> xf_at(V2<double> c, double a, V2<double> lc) {
  XF<double> t;
  t.q = rot_of<double>(a);
  t.p = c - rmul(t.q, lc);
  return t;
}

__device__ inline XF<double> sweep_xf(const TSweep& s, double beta) {  // Sweep.get_transform
  const V2<double> p = scale(1.0 - beta, s.c0) + scale(beta, s.c);
  return xf_at(p, (1.0 - beta) * s.a0 + beta * s.a, s.lc);
}

__device__ inline void sweep_advance(TSweep& s, double alpha) {  // Sweep.advance
  const double beta = (alpha - s.alpha0) / (1.0 - s.alpha0);
  s.c0 = scale(1.0 - beta, s.c0) + scale(beta, s.c);
  s.a0 = (1.0 - beta) * s.a0 + beta * s.a;
  s.alpha0 = alpha;
}

enum : int { SEP_POINTS = 0, SEP_FACE_A, SEP_FACE_B };
struct SepFn {
  const TProxy* pa;
  const TProxy* pb;
  TSweep sa, sb;
  int type;
  V2<double> axis, lp;
};

__device__ inline void sep_init(SepFn& f, const GjkCache& cache, double t1) {  // toi.py:54-122
  const XF<double> xfa = sweep_xf(f.sa, t1), xfb = sweep_xf(f.sb, t1);
  const int ia0 = cache.ia[0], ib0 = cache.ib[0];
  if (cache.count == 1) {
    f.type = SEP_POINTS;
    f.axis = xmul(xfb, f.pb->v[ib0]) - xmul(xfa, f.pa->v[ia0]);  // left unnormalised, as in the reference
    return;
  }
  const int ia1 = cache.ia[1], ib1 = cache.ib[1];
  if (ia0 == ia1) {
    f.type = SEP_FACE_B;
    const V2<double> b1 = f.pb->v[ib0], b2 = f.pb->v[ib1];
    f.axis = normalized(cross_vs(b2 - b1, 1.0));
    const V2<double> n = rmul(xfb.q, f.axis);
    f.lp = scale(0.5, b1 + b2);
    const V2<double> pb = xmul(xfb, f.lp), pa = xmul(xfa, f.pa->v[ia0]);
    if (dot(pa - pb, n) < 0.0) f.axis = -f.axis;
  } else {
    f.type = SEP_FACE_A;
    const V2<double> a1 = f.pa->v[ia0], a2 = f.pa->v[ia1];
    f.axis = normalized(cross_vs(a2 - a1, 1.0));
    const V2<double> n = rmul(xfa.q, f.axis);
    f.lp = scale(0.5, a1 + a2);
    const V2<double> pa = xmul(xfa, f.lp), pb = xmul(xfb, f.pb->v[ib0]);
    if (dot(pb - pa, n) < 0.0) f.axis = -f.axis;
  }
}

// find_min_separation (find = true; toi.py:124-175) or evaluate (toi.py:177-213)
__device__ inline double sep_eval(const SepFn& f, bool find, int& ia, int& ib, double t) {
  const XF<double> xfa = sweep_xf(f.sa, t), xfb = sweep_xf(f.sb, t);
  if (f.type == SEP_POINTS) {
    if (find) {
      ia = tproxy_support(*f.pa, rmul_t(xfa.q, f.axis));
      ib = tproxy_support(*f.pb, rmul_t(xfb.q, -f.axis));
    }
    return dot(xmul(xfb, f.pb->v[ib]) - xmul(xfa, f.pa->v[ia]), f.axis);
  }
  if (f.type == SEP_FACE_A) {
    const V2<double> n = rmul(xfa.q, f.axis), pa = xmul(xfa, f.lp);
    if (find) {
      ia = -1;
      ib = tproxy_support(*f.pb, rmul_t(xfb.q, -n));
    }
    return dot(xmul(xfb, f.pb->v[ib]) - pa, n);
  }
  const V2<double> n = rmul(xfb.q, f.axis), pb = xmul(xfb, f.lp);
  if (find) {
    ib = -1;
    ia = tproxy_support(*f.pa, rmul_t(xfa.q, -n));
  }
  return dot(xmul(xfa, f.pa->v[ia]) - pb, n);
}

// time_of_impact over [0, 1] (toi.py:215-391); returns the state and the time in t_out.
__device__ inline int time_of_impact(const TProxy& pa, const TProxy& pb, const TSweep& swa, const TSweep& swb,
                                     double& t_out) {
  const double t_max = 1.0;
  SepFn f;
  f.pa = &pa;
  f.pb = &pb;
  f.sa = swa;
  f.sb = swb;
  double d = 2.0 * kToiPi * floor(f.sa.a0 / (2.0 * kToiPi));  // Sweep.normalize on the copies
  f.sa.a0 -= d;
  f.sa.a -= d;
  d = 2.0 * kToiPi * floor(f.sb.a0 / (2.0 * kToiPi));
  f.sb.a0 -= d;
  f.sb.a -= d;
  const double total_radius = pa.radius + pb.radius;
  const double target = pmax(kToiLinearSlop, total_radius - 3.0 * kToiLinearSlop);
  const double tolerance = 0.25 * kToiLinearSlop;
  double t1 = 0.0, out_t = t_max;
  int iter = 0, state = TOI_UNKNOWN;
  GjkCache cache;
  cache.count = 0;
  cache.metric = 0.0;
  for (;;) {
    const XF<double> xfa = sweep_xf(f.sa, t1), xfb = sweep_xf(f.sb, t1);
    const double dist = gjk_distance(pa, pb, xfa, xfb, cache);
    if (dist <= 0.0) {
      state = TOI_OVERLAPPED;
      out_t = 0.0;
      break;
    }
    if (dist < target + tolerance) {
      state = TOI_TOUCHING;
      out_t = t1;
      break;
    }
    sep_init(f, cache, t1);
    bool done = false;
    int push_back = 0;
    double t2 = t_max;
    for (;;) {
      int ia = 0, ib = 0;
      double s2 = sep_eval(f, true, ia, ib, t2);
      if (s2 > target + tolerance) {
        state = TOI_SEPARATED;
        out_t = t_max;
        done = true;
        break;
      }
      if (s2 > target - tolerance) {
        t1 = t2;
        break;
      }
      double s1 = sep_eval(f, false, ia, ib, t1);
      if (s1 < target - tolerance) {
        state = TOI_FAILED;
        out_t = t1;
        done = true;
        break;
      }
      if (s1 <= target + tolerance) {
        state = TOI_TOUCHING;
        out_t = t1;
        done = true;
        break;
      }
      int rc = 0;
      double a1 = t1, a2 = t2;
      for (;;) {
        const double t = (rc & 1) ? a1 + (target - s1) * (a2 - a1) / (s2 - s1) : 0.5 * (a1 + a2);
        const double sv = sep_eval(f, false, ia, ib, t);
        if (fabs(sv - target) < tolerance) {
          t2 = t;
          break;
        }
        if (sv > target) {
          a1 = t;
          s1 = sv;
        } else {
          a2 = t;
          s2 = sv;
        }
        ++rc;
        if (rc == 50) break;
      }
      if (++push_back == kToiPushBackLimit) {
        state = TOI_STUCK;
        out_t = t1;
        done = true;
        break;
      }
    }
    ++iter;
    if (done) break;
    if (iter == kToiMaxIters) {
      state = TOI_FAILED;
      out_t = t1;
      break;
    }
  }
  t_out = out_t;
  return state;
}

}  // namespace b2g
