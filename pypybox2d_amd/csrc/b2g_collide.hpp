// b2g_collide.hpp -- narrow-phase manifold generation (one lane per contact).
//
// Restates pypybox2d/collision.py and collision_util.py for the GPU:
//   collide_circles         collision.py:46-66
//   collide_polygon_circle  collision.py:68-141
//   collide_polygons        collision.py:144-241 (+ collision_util.py:31-164)
//   collide_edge_circle     collision.py:246-353
//   collide_edge_polygon    collision.py:355-690 (EPCollider)
// with the reference's tie rules (first index of min/max), hill-climbing SAT
// started from the stored polygon centroid (quirk Q2), hysteresis constants,
// clip ordering and contact ids (quirk Q9 for edge/polygon FACE_B), and the
// untransformed translation of Transform.mul_t(Transform) (quirk Q4).
//
// Shape geometry is read from a per-fixture row in global memory (stride
// GEO_STRIDE scalars), so no per-lane arrays are dynamically indexed.
// Transforms are given in a frame shifted to the position of body A (the
// shift is taken in FP64 before conversion), which keeps FP32 manifolds
// accurate at large world coordinates; all outputs are body-local and do not
// depend on the shift.
#pragma once

#include "b2g_math.hpp"

namespace b2g {

enum { GEO_VERTS = 0, GEO_NORMALS = 16, GEO_CENTROID = 32, GEO_CIRCLE_P = 34, GEO_EDGE = 36, GEO_RADIUS = 44,
       GEO_STRIDE = 48 };
enum { SHAPE_CIRCLE = 0, SHAPE_POLYGON = 1, SHAPE_LOOP = 2, SHAPE_EDGE = 3 };
enum { M_CIRCLES = 0, M_FACE_A = 1, M_FACE_B = 2 };

B2G_HD uint32_t pack_id(int a, int b, int c, int d) {
  return (uint32_t)(a & 255) | ((uint32_t)(b & 255) << 8) | ((uint32_t)(c & 255) << 16) | ((uint32_t)(d & 255) << 24);
}
B2G_HD int id_part(uint32_t id, int k) { return (int)((id >> (8 * k)) & 255u); }

// Shape geometry stored as S and evaluated in T (the FP32 pipeline stores
// FP32 geometry and evaluates it in FP64).
template <typename T, typename S = T>
struct Geo {
  const S* g;
  int n, type, has0, has3;
  B2G_HD V2<T> v(int i) const { return v2<T>((T)g[GEO_VERTS + 2 * i], (T)g[GEO_VERTS + 2 * i + 1]); }
  B2G_HD V2<T> nrm(int i) const { return v2<T>((T)g[GEO_NORMALS + 2 * i], (T)g[GEO_NORMALS + 2 * i + 1]); }
  B2G_HD V2<T> centroid() const { return v2<T>((T)g[GEO_CENTROID], (T)g[GEO_CENTROID + 1]); }
  B2G_HD V2<T> cp() const { return v2<T>((T)g[GEO_CIRCLE_P], (T)g[GEO_CIRCLE_P + 1]); }
  B2G_HD V2<T> e(int k) const { return v2<T>((T)g[GEO_EDGE + 2 * k], (T)g[GEO_EDGE + 2 * k + 1]); }
  B2G_HD T radius() const { return (T)g[GEO_RADIUS]; }
};

template <typename T>
struct Man {
  int type, pc;
  V2<T> ln, lp, p0, p1;
  uint32_t id0, id1;
};

template <typename T>
struct ClipV {
  V2<T> v;
  uint32_t id;
};

template <typename T>
struct Consts {
  static constexpr T MAX_FLOAT = (T)3.0e38;
  static constexpr T EPS = NormEps<T>::value;
  static constexpr T LINEAR_SLOP = (T)0.005;
  static constexpr T POLYGON_RADIUS = (T)(2.0 * 0.005);
  static constexpr T ANGULAR_SLOP = (T)(2.0 / 180.0 * 3.141592653589793);
  static constexpr T RELATIVE_TOL = (T)0.98;
  static constexpr T ABSOLUTE_TOL = (T)0.001;
};

template <typename T, typename S>
B2G_HD void collide_circles(Man<T>& m, const Geo<T, S>& A, XF<T> xfa, const Geo<T, S>& B, XF<T> xfb) {
  m.pc = 0;
  V2<T> pa = xmul(xfa, A.cp()), pb = xmul(xfb, B.cp());
  V2<T> d = pb - pa;
  T dist_sqr = sq(d.x) + sq(d.y);
  T radius = A.radius() + B.radius();
  if (dist_sqr > sq(radius)) return;
  m.type = M_CIRCLES;
  m.lp = A.cp();
  m.ln = v2<T>(0, 0);
  m.pc = 1;
  m.p0 = B.cp();
  m.id0 = 0;
}

template <typename T, typename S>
B2G_HD void collide_polygon_circle(Man<T>& m, const Geo<T, S>& P, XF<T> xfa, const Geo<T, S>& C, XF<T> xfb) {
  m.pc = 0;
  V2<T> c = xmul(xfb, C.cp());
  V2<T> cl = xmul_t(xfa, c);
  int normal_index = 0;
  T separation = -Consts<T>::MAX_FLOAT;
  T radius = P.radius() + C.radius();
  const int n = P.n;
  for (int i = 0; i < n; ++i) {
    T s = dot(P.nrm(i), cl - P.v(i));
    if (s > radius) return;
    if (s > separation) {
      separation = s;
      normal_index = i;
    }
  }
  int i1 = normal_index, i2 = (i1 + 1) % n;
  V2<T> v1 = P.v(i1), vv2 = P.v(i2);
  if (separation < Consts<T>::EPS) {
    m.pc = 1;
    m.type = M_FACE_A;
    m.ln = P.nrm(normal_index);
    m.lp = scale((T)0.5, v1 + vv2);
    m.p0 = C.cp();
    m.id0 = 0;
    return;
  }
  T u1 = dot(cl - v1, vv2 - v1);
  T u2 = dot(cl - vv2, v1 - vv2);
  if (u1 < (T)0) {
    V2<T> d = cl - v1;
    if (dot(d, d) > sq(radius)) return;
    m.ln = normalized(cl - v1);
    m.lp = v1;
  } else if (u2 < (T)0) {
    V2<T> d = cl - vv2;
    if (dot(d, d) > sq(radius)) return;
    m.ln = normalized(cl - vv2);
    m.lp = vv2;
  } else {
    V2<T> fc = scale((T)0.5, v1 + vv2);
    separation = dot(cl - fc, P.nrm(i1));
    if (separation > radius) return;
    m.ln = P.nrm(i1);
    m.lp = fc;
  }
  m.pc = 1;
  m.type = M_FACE_A;
  m.p0 = C.cp();
  m.id0 = 0;
}

// collision_util.py:31-52
template <typename T, typename S>
B2G_HD T edge_separation(const Geo<T, S>& p1, XF<T> xf1, int edge1, const Geo<T, S>& p2, XF<T> xf2) {
  V2<T> n1w = rmul(xf1.q, p1.nrm(edge1));
  V2<T> n1 = rmul_t(xf2.q, n1w);
  int index = 0;
  T mind = dot(p2.v(0), n1);
  for (int i = 1; i < p2.n; ++i) {
    T d = dot(p2.v(i), n1);
    if (d < mind) {
      mind = d;
      index = i;
    }
  }
  V2<T> va = xmul(xf1, p1.v(edge1));
  V2<T> vb = xmul(xf2, p2.v(index));
  return dot(vb - va, n1w);
}

// collision_util.py:54-109
template <typename T, typename S>
B2G_HD T max_separation(const Geo<T, S>& p1, XF<T> xf1, const Geo<T, S>& p2, XF<T> xf2, int& edge_out) {
  const int count1 = p1.n;
  V2<T> d = xmul(xf2, p2.centroid()) - xmul(xf1, p1.centroid());
  V2<T> dl = rmul_t(xf1.q, d);
  int edge = 0;
  T maxd = dot(p1.nrm(0), dl);
  for (int i = 1; i < count1; ++i) {
    T v = dot(p1.nrm(i), dl);
    if (v > maxd) {
      maxd = v;
      edge = i;
    }
  }
  T s = edge_separation(p1, xf1, edge, p2, xf2);
  int prev_edge = pymod(edge - 1, count1);
  T s_prev = edge_separation(p1, xf1, prev_edge, p2, xf2);
  int next_edge = pymod(edge + 1, count1);
  T s_next = edge_separation(p1, xf1, next_edge, p2, xf2);
  int increment, best_edge;
  T best_sep;
  if (s_prev > s && s_prev > s_next) {
    increment = -1;
    best_edge = prev_edge;
    best_sep = s_prev;
  } else if (s_next > s) {
    increment = 1;
    best_edge = next_edge;
    best_sep = s_next;
  } else {
    edge_out = edge;
    return s;
  }
  for (int guard = 0; guard < count1; ++guard) {
    edge = increment == -1 ? pymod(best_edge - 1, count1) : pymod(best_edge + 1, count1);
    s = edge_separation(p1, xf1, edge, p2, xf2);
    if (s > best_sep) {
      best_edge = edge;
      best_sep = s;
    } else {
      break;
    }
  }
  edge_out = best_edge;
  return best_sep;
}

// collision_util.py:111-133
template <typename T, typename S>
B2G_HD void find_incident_edge(ClipV<T>& c0, ClipV<T>& c1, const Geo<T, S>& p1, XF<T> xf1, int edge1, const Geo<T, S>& p2,
                               XF<T> xf2) {
  V2<T> n1 = rmul_t(xf2.q, rmul(xf1.q, p1.nrm(edge1)));
  int i1 = 0;
  T mind = dot(n1, p2.nrm(0));
  for (int i = 1; i < p2.n; ++i) {
    T d = dot(n1, p2.nrm(i));
    if (d < mind) {
      mind = d;
      i1 = i;
    }
  }
  int i2 = (i1 + 1) % p2.n;
  c0.v = xmul(xf2, p2.v(i1));
  c0.id = pack_id(edge1, i1, 1, 0);
  c1.v = xmul(xf2, p2.v(i2));
  c1.id = pack_id(edge1, i2, 1, 0);
}

// collision_util.py:135-164 (Sutherland-Hodgman); at most two points come out.
template <typename T>
B2G_HD int clip_segment_to_line(ClipV<T>& o0, ClipV<T>& o1, ClipV<T> i0, ClipV<T> i1, V2<T> normal, T offset,
                                int via) {
  T d0 = dot(normal, i0.v) - offset;
  T d1 = dot(normal, i1.v) - offset;
  int n = 0;
  if (d0 <= (T)0) {
    o0 = i0;
    n = 1;
  }
  if (d1 <= (T)0) {
    if (n == 0) o0 = i1;
    else o1 = i1;
    ++n;
  }
  if (d0 * d1 < (T)0) {
    ClipV<T> c;
    T interp = d0 / (d0 - d1);
    c.v = i0.v + scale(interp, i1.v - i0.v);
    c.id = pack_id(via, id_part(i0.id, 1), 0, 1);
    if (n == 0) o0 = c;
    else o1 = c;
    ++n;
  }
  return n;
}

template <typename T, typename S>
B2G_HD void collide_polygons(Man<T>& m, const Geo<T, S>& A, XF<T> xfa, const Geo<T, S>& B, XF<T> xfb) {
  m.pc = 0;
  T total_radius = A.radius() + B.radius();
  int edge_a, edge_b;
  T sa = max_separation(A, xfa, B, xfb, edge_a);
  if (sa > total_radius) return;
  T sb = max_separation(B, xfb, A, xfa, edge_b);
  if (sb > total_radius) return;
  const bool flip = sb > Consts<T>::RELATIVE_TOL * sa + Consts<T>::ABSOLUTE_TOL;
  const Geo<T, S>& p1 = flip ? B : A;
  const Geo<T, S>& p2 = flip ? A : B;
  XF<T> xf1 = flip ? xfb : xfa;
  XF<T> xf2 = flip ? xfa : xfb;
  int edge1 = flip ? edge_b : edge_a;
  m.type = flip ? M_FACE_B : M_FACE_A;
  ClipV<T> i0, i1;
  find_incident_edge(i0, i1, p1, xf1, edge1, p2, xf2);
  int iv1 = edge1, iv2 = (edge1 + 1) % p1.n;
  V2<T> v11 = p1.v(iv1), v12 = p1.v(iv2);
  V2<T> ltan = normalized(v12 - v11);
  V2<T> local_normal = cross_vs(ltan, (T)1);
  V2<T> planepoint = scale((T)0.5, v11 + v12);
  V2<T> tangent = rmul(xf1.q, ltan);
  V2<T> normal = cross_vs(tangent, (T)1);
  v11 = xmul(xf1, v11);
  v12 = xmul(xf1, v12);
  T front_offset = dot(normal, v11);
  T side1 = -dot(tangent, v11) + total_radius;
  T side2 = dot(tangent, v12) + total_radius;
  ClipV<T> a0, a1, b0, b1;
  if (clip_segment_to_line(a0, a1, i0, i1, -tangent, side1, iv1) < 2) return;
  if (clip_segment_to_line(b0, b1, a0, a1, tangent, side2, iv2) < 2) return;
  m.ln = local_normal;
  m.lp = planepoint;
  int pc = 0;
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    ClipV<T> cp = i == 0 ? b0 : b1;
    T separation = dot(normal, cp.v) - front_offset;
    if (separation <= total_radius) {
      V2<T> lp = xmul_t(xf2, cp.v);
      uint32_t id = cp.id;
      if (flip) id = pack_id(id_part(id, 1), id_part(id, 0), id_part(id, 3), id_part(id, 2));
      if (pc == 0) {
        m.p0 = lp;
        m.id0 = id;
      } else {
        m.p1 = lp;
        m.id1 = id;
      }
      ++pc;
    }
  }
  m.pc = pc;
}

template <typename T, typename S>
B2G_HD void collide_edge_circle(Man<T>& m, const Geo<T, S>& E, XF<T> xfa, const Geo<T, S>& C, XF<T> xfb) {
  m.pc = 0;
  V2<T> Q = xmul_t(xfa, xmul(xfb, C.cp()));
  V2<T> A = E.e(1), B = E.e(2);
  V2<T> e = B - A;
  T u = dot(e, B - Q);
  T v = dot(e, Q - A);
  T radius = E.radius() + C.radius();
  if (v <= (T)0) {
    V2<T> d = Q - A;
    if (dot(d, d) > sq(radius)) return;
    if (E.has0) {
      V2<T> A1 = E.e(0), B1 = A;
      V2<T> e1 = B1 - A1;
      if (dot(e1, B1 - Q) > (T)0) return;
    }
    m.pc = 1;
    m.type = M_CIRCLES;
    m.ln = v2<T>(0, 0);
    m.lp = A;
    m.id0 = pack_id(0, 0, 0, 0);
    m.p0 = C.cp();
    return;
  }
  if (u <= (T)0) {
    V2<T> d = Q - B;
    if (dot(d, d) > sq(radius)) return;
    if (E.has3) {
      V2<T> B2 = E.e(3), A2 = B;
      V2<T> e2 = B2 - A2;
      if (dot(e2, Q - A2) > (T)0) return;
    }
    m.pc = 1;
    m.type = M_CIRCLES;
    m.ln = v2<T>(0, 0);
    m.lp = B;
    m.id0 = pack_id(1, 0, 0, 0);
    m.p0 = C.cp();
    return;
  }
  T den = dot(e, e);
  V2<T> P = scale((T)1 / den, scale(u, A) + scale(v, B));
  V2<T> d = Q - P;
  if (dot(d, d) > sq(radius)) return;
  V2<T> n = v2<T>(-e.y, e.x);
  if (dot(n, Q - A) < (T)0) n = v2<T>(-n.x, -n.y);
  n = normalized(n);
  m.pc = 1;
  m.type = M_FACE_A;
  m.ln = n;
  m.lp = A;
  m.id0 = pack_id(0, 0, 1, 0);
  m.p0 = C.cp();
}

// ---- EPCollider (collision.py:355-690) ----
template <typename T, typename S>
struct EPC {
  XFM<T> xf;
  Geo<T, S> P;
  V2<T> v0, v1, v2, v3, normal;
  T radius;
  V2<T> l11, l12, l21, l22;
  // proxy B (the polygon moved into the edge frame), computed on demand
  B2G_HD V2<T> bv(int i) const { return xmul(xf, P.v(i)); }
  B2G_HD V2<T> bn(int i) const { return mmul(xf.R, P.nrm(i)); }
  // proxy A (the edge, two-sided)
  B2G_HD V2<T> av(int i) const { return i == 0 ? v1 : v2; }
  B2G_HD V2<T> an(int i) const { return i == 0 ? normal : -normal; }
  B2G_HD bool valid(V2<T> n) const {
    const T s = -Consts<T>::ANGULAR_SLOP;
    bool valid1 = (cross(n, l11) >= s) && (cross(l12, n) >= s);
    bool valid2 = (cross(n, l21) >= s) && (cross(l22, n) >= s);
    return valid1 && valid2;
  }
};

template <typename T, typename S>
B2G_HD void ep_adjacency(EPC<T, S>& c, bool has0, bool has3) {  // collision.py:541-604
  V2<T> cb = xmul(c.xf, c.P.centroid());
  if (has0) {
    V2<T> e0 = c.v1 - c.v0, e1 = c.v2 - c.v1;
    V2<T> n0 = normalized(v2<T>(e0.y, -e0.x)), n1 = normalized(v2<T>(e1.y, -e1.x));
    bool convex = cross(n0, n1) >= (T)0;
    bool front0 = dot(n0, cb - c.v0) >= (T)0;
    bool front1 = dot(n1, cb - c.v1) >= (T)0;
    if (convex) {
      if (front0 || front1) { c.l11 = n1; c.l12 = n0; }
      else { c.l11 = -n1; c.l12 = -n0; }
    } else {
      if (front0 && front1) { c.l11 = n0; c.l12 = n1; }
      else { c.l11 = -n0; c.l12 = -n1; }
    }
  } else {
    c.l11 = v2<T>(0, 0);
    c.l12 = v2<T>(0, 0);
  }
  if (has3) {
    V2<T> e1 = c.v2 - c.v1, e2 = c.v3 - c.v2;
    V2<T> n1 = normalized(v2<T>(e1.y, -e1.x)), n2 = normalized(v2<T>(e2.y, -e2.x));
    bool convex = cross(n1, n2) >= (T)0;
    bool front1 = dot(n1, cb - c.v1) >= (T)0;
    bool front2 = dot(n2, cb - c.v2) >= (T)0;
    if (convex) {
      if (front1 || front2) { c.l21 = n2; c.l22 = n1; }
      else { c.l21 = -n2; c.l22 = -n1; }
    } else {
      if (front1 && front2) { c.l21 = n1; c.l22 = n2; }
      else { c.l21 = -n1; c.l22 = -n2; }
    }
  } else {
    c.l21 = v2<T>(0, 0);
    c.l22 = v2<T>(0, 0);
  }
}

enum { EP_UNKNOWN = 0, EP_EDGE_A = 1, EP_EDGE_B = 2 };

template <typename T, typename S>
B2G_HD void collide_edge_polygon(Man<T>& m, const Geo<T, S>& E, XF<T> xfa, const Geo<T, S>& P, XF<T> xfb) {
  EPC<T, S> c;
  c.xf = xmul_t_xf(xfa, xfb);
  c.P = P;
  c.v0 = E.e(0);
  c.v1 = E.e(1);
  c.v2 = E.e(2);
  c.v3 = E.e(3);
  V2<T> e = c.v2 - c.v1;
  c.normal = normalized(v2<T>(e.y, -e.x));
  c.radius = (T)2 * Consts<T>::POLYGON_RADIUS;
  m.pc = 0;
  ep_adjacency(c, E.has0 != 0, E.has3 != 0);
  // compute_edge_separation (collision.py:606-635)
  int ea_type = EP_UNKNOWN, ea_index = -1;
  T ea_sep = -Consts<T>::MAX_FLOAT;
  for (int i = 0; i < 2; ++i) {
    V2<T> n = c.an(i);
    if (!c.valid(n)) continue;
    T sep = dot(n, c.bv(0) - c.v1);
    for (int k = 1; k < P.n; ++k) {
      T d = dot(n, c.bv(k) - c.v1);
      if (d < sep) sep = d;
    }
    if (sep > c.radius) {
      ea_type = EP_EDGE_A;
      ea_index = i;
      ea_sep = sep;
      break;
    }
    if (sep > ea_sep) {
      ea_type = EP_EDGE_A;
      ea_index = i;
      ea_sep = sep;
    }
  }
  if (ea_type == EP_UNKNOWN) return;
  if (ea_sep > c.radius) return;
  // compute_polygon_separation (collision.py:637-666)
  int pa_type = EP_UNKNOWN, pa_index = -1;
  T pa_sep = -Consts<T>::MAX_FLOAT;
  for (int i = 0; i < P.n; ++i) {
    V2<T> n = -c.bn(i);
    if (!c.valid(n)) continue;
    V2<T> vb = c.bv(i);
    T s1 = dot(n, vb - c.v1);
    T s2 = dot(n, vb - c.v2);
    T s = pmin(s1, s2);
    if (s > c.radius) {
      pa_type = EP_EDGE_B;
      pa_index = i;
      pa_sep = s;
    }
    if (s > pa_sep) {
      pa_type = EP_EDGE_B;
      pa_index = i;
      pa_sep = s;
    }
  }
  if (pa_type != EP_UNKNOWN && pa_sep > c.radius) return;
  bool primary_b;
  if (pa_type == EP_UNKNOWN) primary_b = false;
  else if (pa_sep > Consts<T>::RELATIVE_TOL * ea_sep + Consts<T>::ABSOLUTE_TOL) primary_b = true;
  else primary_b = false;
  const int edge1 = primary_b ? pa_index : ea_index;
  m.type = primary_b ? M_FACE_B : M_FACE_A;
  // find_incident_edge (collision.py:668-690): proxy1 = primary side
  const int n1cnt = primary_b ? P.n : 2;
  const int n2cnt = primary_b ? 2 : P.n;
  V2<T> rn = primary_b ? c.bn(edge1) : c.an(edge1);
  int i1 = 0;
  T mind = dot(rn, primary_b ? c.an(0) : c.bn(0));
  for (int i = 1; i < n2cnt; ++i) {
    T d = dot(rn, primary_b ? c.an(i) : c.bn(i));
    if (d < mind) {
      mind = d;
      i1 = i;
    }
  }
  int i2 = (i1 + 1) % n2cnt;
  ClipV<T> in0, in1;
  in0.v = primary_b ? c.av(i1) : c.bv(i1);
  in0.id = pack_id(edge1, i1, 1, 0);
  in1.v = primary_b ? c.av(i2) : c.bv(i2);
  in1.id = pack_id(edge1, i2, 1, 0);
  int iv1 = edge1, iv2 = (edge1 + 1) % n1cnt;
  V2<T> v11 = primary_b ? c.bv(iv1) : c.av(iv1);
  V2<T> v12 = primary_b ? c.bv(iv2) : c.av(iv2);
  V2<T> tangent = normalized(v12 - v11);
  V2<T> normal = cross_vs(tangent, (T)1);
  V2<T> plane_point = scale((T)0.5, v11 + v12);
  T front_offset = dot(normal, v11);
  T side1 = -dot(tangent, v11) + c.radius;
  T side2 = dot(tangent, v12) + c.radius;
  ClipV<T> a0, a1, b0, b1;
  if (clip_segment_to_line(a0, a1, in0, in1, -tangent, side1, iv1) < 2) return;
  int n2 = clip_segment_to_line(b0, b1, a0, a1, tangent, side2, iv2);
  if (n2 < 2) return;
  if (!primary_b) {
    m.ln = normal;
    m.lp = plane_point;
  } else {
    m.ln = mmul_t(c.xf.R, normal);
    m.lp = xmul_t(c.xf, plane_point);
  }
  int pc = 0;
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    ClipV<T> cp = i == 0 ? b0 : b1;
    T separation = dot(normal, cp.v) - front_offset;
    if (separation <= c.radius) {
      V2<T> lp;
      uint32_t id;
      if (!primary_b) {
        lp = xmul_t(c.xf, cp.v);
        id = cp.id;
      } else {
        lp = cp.v;
        id = pack_id(id_part(cp.id, 3), id_part(cp.id, 2), id_part(cp.id, 1), id_part(cp.id, 0));
      }
      if (pc == 0) {
        m.p0 = lp;
        m.id0 = id;
      } else {
        m.p1 = lp;
        m.id1 = id;
      }
      ++pc;
    }
  }
  m.pc = pc;
}

// Contact.evaluate dispatch for registry-oriented fixtures (contact.py:343-373).
template <typename T, typename S>
B2G_HD bool evaluate(Man<T>& m, const Geo<T, S>& A, XF<T> xfa, const Geo<T, S>& B, XF<T> xfb) {
  const int ta = A.type, tb = B.type;
  if (ta == SHAPE_CIRCLE && tb == SHAPE_CIRCLE) collide_circles(m, A, xfa, B, xfb);
  else if (ta == SHAPE_POLYGON && tb == SHAPE_POLYGON) collide_polygons(m, A, xfa, B, xfb);
  else if (ta == SHAPE_POLYGON && tb == SHAPE_CIRCLE) collide_polygon_circle(m, A, xfa, B, xfb);
  else if (ta == SHAPE_EDGE && tb == SHAPE_CIRCLE) collide_edge_circle(m, A, xfa, B, xfb);
  else if (ta == SHAPE_EDGE && tb == SHAPE_POLYGON) collide_edge_polygon(m, A, xfa, B, xfb);
  else return false;
  return true;
}

}  // namespace b2g
