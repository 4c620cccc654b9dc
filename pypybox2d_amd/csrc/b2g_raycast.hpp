// b2g_raycast.hpp -- ray casts of one shape and the leaf test of the tree ray cast, FP64.
//
// Shape ray casts follow pypybox2d's shapes.py operation by operation (binary64, no
// FMA contraction: the library is built with -ffp-contract=off):
//   Circle.ray_cast   shapes.py:147-174
//   Polygon.ray_cast  shapes.py:265-307
//   Edge.ray_cast     shapes.py:655-702 (the normal is returned in the edge's
//                     frame, not rotated to world coordinates, as the reference does)
// The ray is p1 -> p2 with max_fraction 1 (World.ray_cast, world.py:426-440,
// always asks the fixture with 1.0).  The leaf test is DynamicTree.ray_cast's
// (dynamictree.py:209-273): the segment's bounding box against the proxy's fat
// AABB (inclusive overlap), then the separating axis |v.(p1 - c)| - |v|.h > 0.
#pragma once

#include "b2g_collide.hpp"

namespace b2g {

struct RayHit {
  double fraction, nx, ny;
};

// Transform of a body: position and rotation (c, s) (col1 = (c, s), col2 = (-s, c)).
struct XF64 {
  double px, py, c, s;
};

// R^T (v - p) (PyTransform.mul_t with a vector, common.py:662-670; PyMat22.mul_t, 474-490)
__device__ __forceinline__ V2<double> to_local(const XF64& x, double vx, double vy) {
  const double dx = vx - x.px, dy = vy - x.py;
  return v2<double>(x.c * dx + x.s * dy, (-x.s) * dx + x.c * dy);
}

// Vec2.length and normalized with the pure-Python flavour's threshold (common.py:186-222)
__device__ __forceinline__ V2<double> py_normalized(V2<double> a) {
  const double l = sqrt(a.x * a.x + a.y * a.y);
  if (l < NormEps<double>::value) return a;
  const double il = 1.0 / l;
  return v2<double>(a.x * il, a.y * il);
}

__device__ inline bool ray_circle(const Geo<double, double>& g, const XF64& x, double p1x, double p1y, double p2x,
                                  double p2y, RayHit& h) {
  const V2<double> cp = g.cp();
  // position = xf.position + xf.rotation * self.position
  const double qx = x.px + (x.c * cp.x + (-x.s) * cp.y), qy = x.py + (x.s * cp.x + x.c * cp.y);
  const double sx = p1x - qx, sy = p1y - qy;
  const double r = g.radius();
  const double b = (sx * sx + sy * sy) - r * r;
  const double rx = p2x - p1x, ry = p2y - p1y;
  const double c = sx * rx + sy * ry;
  const double rr = rx * rx + ry * ry;
  const double sigma = c * c - rr * b;
  if (sigma < 0.0 || rr < NormEps<double>::value) return false;
  double a = -(c + sqrt(sigma));
  if (0.0 <= a && a <= 1.0 * rr) {
    a /= rr;
    h.fraction = a;
    const V2<double> n = py_normalized(v2<double>(sx + a * rx, sy + a * ry));
    h.nx = n.x;
    h.ny = n.y;
    return true;
  }
  return false;
}

__device__ inline bool ray_polygon(const Geo<double, double>& g, const XF64& x, double p1x, double p1y, double p2x,
                                   double p2y, RayHit& h) {
  const V2<double> a = to_local(x, p1x, p1y), b = to_local(x, p2x, p2y);
  const double dx = b.x - a.x, dy = b.y - a.y;
  double lower = 0.0, upper = 1.0;
  int tagged = -1;
  for (int i = 0; i < g.n; ++i) {
    const V2<double> nrm = g.nrm(i), v = g.v(i);
    const double numerator = nrm.x * (v.x - a.x) + nrm.y * (v.y - a.y);
    const double denominator = nrm.x * dx + nrm.y * dy;
    if (denominator == 0.0) {
      if (numerator < 0.0) return false;
    } else if (denominator < 0.0 && numerator < lower * denominator) {
      lower = numerator / denominator;
      tagged = i;
    } else if (denominator > 0.0 && numerator < upper * denominator) {
      upper = numerator / denominator;
    }
    if (upper < lower) return false;
  }
  if (tagged < 0) return false;  // the ray starts inside
  const V2<double> n = g.nrm(tagged);
  h.nx = x.c * n.x + (-x.s) * n.y;
  h.ny = x.s * n.x + x.c * n.y;
  h.fraction = lower;
  return true;
}

__device__ inline bool ray_edge(const Geo<double, double>& g, const XF64& x, double p1x, double p1y, double p2x,
                                double p2y, RayHit& h) {
  const V2<double> a = to_local(x, p1x, p1y), b = to_local(x, p2x, p2y);
  const double dx = b.x - a.x, dy = b.y - a.y;
  const V2<double> v1 = g.e(1), v2_ = g.e(2);
  const double ex = v2_.x - v1.x, ey = v2_.y - v1.y;
  V2<double> nrm = py_normalized(v2<double>(ey, -ex));
  const double numerator = nrm.x * (v1.x - a.x) + nrm.y * (v1.y - a.y);
  const double denominator = nrm.x * dx + nrm.y * dy;
  if (denominator == 0.0) return false;
  const double t = numerator / denominator;
  if (t < 0.0 || 1.0 < t) return false;
  const double qx = a.x + t * dx, qy = a.y + t * dy;
  const double rx = v2_.x - v1.x, ry = v2_.y - v1.y;
  const double rr = rx * rx + ry * ry;
  if (rr == 0.0) return false;
  const double s = ((qx - v1.x) * rx + (qy - v1.y) * ry) / rr;
  if (s < 0.0 || 1.0 < s) return false;
  h.fraction = t;
  if (numerator > 0.0) nrm = -nrm;
  h.nx = nrm.x;
  h.ny = nrm.y;
  return true;
}

__device__ inline bool ray_shape(const Geo<double, double>& g, const XF64& x, double p1x, double p1y, double p2x,
                                 double p2y, RayHit& h) {
  if (g.type == SHAPE_CIRCLE) return ray_circle(g, x, p1x, p1y, p2x, p2y, h);
  if (g.type == SHAPE_POLYGON) return ray_polygon(g, x, p1x, p1y, p2x, p2y, h);
  if (g.type == SHAPE_EDGE) return ray_edge(g, x, p1x, p1y, p2x, p2y, h);
  return false;
}

// The segment's data for the tree ray cast (dynamictree.py:225-241): v = scalar_cross(1, r) of
// the normalized direction, |v|, and the segment's bounding box for max_fraction 1
// (t = p1 + 1.0 * (p2 - p1)).
struct RaySeg {
  double p1x, p1y, p2x, p2y, vx, vy, avx, avy, lx, ly, ux, uy;
};

__device__ __forceinline__ RaySeg ray_seg(double p1x, double p1y, double p2x, double p2y) {
  RaySeg s;
  s.p1x = p1x; s.p1y = p1y; s.p2x = p2x; s.p2y = p2y;
  const V2<double> r = py_normalized(v2<double>(p2x - p1x, p2y - p1y));
  s.vx = -1.0 * r.y;
  s.vy = 1.0 * r.x;
  s.avx = fabs(s.vx);
  s.avy = fabs(s.vy);
  const double tx = p1x + 1.0 * (p2x - p1x), ty = p1y + 1.0 * (p2y - p1y);
  s.lx = pmin(p1x, tx); s.ly = pmin(p1y, ty);
  s.ux = pmax(p1x, tx); s.uy = pmax(p1y, ty);
  return s;
}

// node.aabb.test_overlap(segment_aabb) (PyAABB.test_overlap: a = node box, b = segment box)
__device__ __forceinline__ bool seg_box_overlap(const RaySeg& s, double lx, double ly, double ux, double uy) {
  return !(s.lx - ux > 0.0 || s.ly - uy > 0.0 || lx - s.ux > 0.0 || ly - s.uy > 0.0);
}

// The leaf test of DynamicTree.ray_cast on a proxy's fat AABB.
__device__ __forceinline__ bool ray_leaf_passes(const RaySeg& s, double lx, double ly, double ux, double uy) {
  if (!seg_box_overlap(s, lx, ly, ux, uy)) return false;
  const double cx = 0.5 * (lx + ux), cy = 0.5 * (ly + uy);
  const double hx = 0.5 * (ux - lx), hy = 0.5 * (uy - ly);
  const double sep = fabs(s.vx * (s.p1x - cx) + s.vy * (s.p1y - cy)) - (s.avx * hx + s.avy * hy);
  return !(sep > 0.0);
}

}  // namespace b2g
