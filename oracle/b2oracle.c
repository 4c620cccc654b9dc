/*
 * b2oracle.c -- FP64 CPU restatement of pypybox2d's per-step contact pipeline.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle the GPU library is
 * checked against; it is loaded by tests/, by __graft_entry__.smoke() as the
 * checker, and by bench.py's cpu_baseline leg.  No product code path links,
 * loads or calls it.
 *
 * It follows the pure-Python flavour of the reference (pypybox2d/common.py
 * PyVec2/PyMat22/PyTransform/PyAABB) operation by operation, so that with the
 * same glibc it reproduces CPython's binary64 results bit for bit:
 *   - `x ** 2` is CPython float_pow -> libm pow(|x|, 2.0) (not x*x),
 *   - `min(a, b)` / `max(a, b)` keep the first argument on ties,
 *   - expressions are evaluated left to right with no FMA contraction
 *     (build with -ffp-contract=off -fno-builtin).
 * Reference quirks reproduced (see the cited lines):
 *   Q2  Polygon._centroid stays as given (it is (0,0) for every polygon the
 *       reference builds, shapes.py:244,482).
 *   Q4  Transform.mul_t(Transform) does not rotate the translation
 *       (common.py:671-673), used by the edge/polygon collider.
 *   Q6  ill-conditioned 2-point block: the velocity constraint keeps only its
 *       first point; the position constraint keeps both (contact.py:487-496).
 *   Q7  sleep test reads the stale loop variable (island.py:163-179).
 *   Q9  edge/polygon FACE_B ids are (type_b, type_a, index_b, index_a)
 *       (collision.py:532).
 * Contact-creation order (Q1): the reference iterates an identity-hashed set
 * (broadphase.py:98-109); the oracle emits new pairs in move-buffer order with
 * ascending partner index, or, when a hint is installed, in the hint's order
 * (used to replay a recorded reference run).
 *
 * The contact listener (begin_contact, end_contact, pre_solve, post_solve)
 * is recorded as event records in call order (orc_events), and a custom
 * contact_filter is given as its verdict for every ordered fixture pair
 * (orc_set_filter).
 *
 * Joints (pypybox2d/joints/) are solved in the island's Gauss-Seidel order
 * (island.py:107-148), their quirks reproduced where the reference has them
 * (cited at each joint type).
 *
 * Not supported (reported through the error word): sensors, loop shapes,
 * controllers.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/b2g_layout.h"

#define ORC_EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------ */
/* settings.py:30-123                                                  */
/* ------------------------------------------------------------------ */
static const double PI_ = 3.141592653589793;
#define EPSILON DBL_EPSILON
#define MAX_FLOAT DBL_MAX
static const double AABB_EXTENSION = 0.1;   /* settings.py:52 */
static const double AABB_MULTIPLIER = 2.0;  /* settings.py:57 */
static const double LINEAR_SLOP = 0.005;    /* settings.py:61 */
static const double VELOCITY_THRESHOLD = 1.0;
static const double MAX_LINEAR_CORRECTION = 0.2;
static const double MAX_TRANSLATION = 2.0;
static const double BAUMGARTE = 0.2;
static const double TIME_TO_SLEEP = 0.5;
static double MAX_ANGULAR_CORRECTION;
static double ANGULAR_SLOP, POLYGON_RADIUS, MAX_ROTATION, MAX_ROTATION_SQR,
    MAX_TRANSLATION_SQR, LINEAR_SLEEP_TOLERANCE_SQR, ANGULAR_SLEEP_TOLERANCE_SQR,
    EPSILON_SQR;

/* CPython float_pow(x, 2) (Objects/floatobject.c): 0 -> 0.0, |x| -> pow */
static double (*volatile pow_fn)(double, double) = pow;
static double pysq(double x) {
  if (x != x) return x;
  if (x == 0.0) return 0.0;
  if (x < 0.0) x = -x;
  if (x == 1.0) return 1.0;
  return pow_fn(x, 2.0);
}

static int g_init = 0;
static void init_consts(void) {
  if (g_init) return;
  ANGULAR_SLOP = 2.0 / 180.0 * PI_;        /* settings.py:65 */
  MAX_ANGULAR_CORRECTION = 8.0 / 180.0 * PI_; /* settings.py:94 */
  POLYGON_RADIUS = 2.0 * LINEAR_SLOP;      /* settings.py:70 */
  MAX_TRANSLATION_SQR = pysq(MAX_TRANSLATION);
  MAX_ROTATION = 0.5 * PI_;                /* settings.py:103 */
  MAX_ROTATION_SQR = pysq(MAX_ROTATION);
  LINEAR_SLEEP_TOLERANCE_SQR = pysq(0.01); /* settings.py:118-119 */
  ANGULAR_SLEEP_TOLERANCE_SQR = pysq(2.0 / 180.0 * PI_);
  EPSILON_SQR = pysq(EPSILON);             /* settings.py:39 */
  g_init = 1;
}

/* ------------------------------------------------------------------ */
/* common.py PyVec2 / PyMat22 / PyTransform / PyAABB                   */
/* ------------------------------------------------------------------ */
typedef struct { double x, y; } V2;
typedef struct { double c1x, c1y, c2x, c2y; } M22; /* column-major, common.py:360 */
typedef struct { V2 p; M22 R; } XF;
typedef struct { V2 lo, hi; } AABB;

static inline V2 v2(double x, double y) { V2 r; r.x = x; r.y = y; return r; }
static inline V2 add(V2 a, V2 b) { return v2(a.x + b.x, a.y + b.y); }
static inline V2 sub(V2 a, V2 b) { return v2(a.x - b.x, a.y - b.y); }
static inline V2 neg(V2 a) { return v2(-a.x, -a.y); }
/* Vec2.__mul__(scalar) = Vec2(value*x, value*y) (common.py:113-117) */
static inline V2 scale(double s, V2 a) { return v2(s * a.x, s * a.y); }
static inline double dot(V2 a, V2 b) { return a.x * b.x + a.y * b.y; }
static inline double cross(V2 a, V2 b) { return a.x * b.y - a.y * b.x; }
/* Vec2.cross(scalar) = (s*y, -s*x) (common.py:200-204) */
static inline V2 cross_vs(V2 a, double s) { return v2(s * a.y, -s * a.x); }
/* scalar_cross(s, v) = v._scalar_cross(s) = (-s*y, s*x) (common.py:188-195) */
static inline V2 cross_sv(double s, V2 a) { return v2(-s * a.y, s * a.x); }
static inline double length_(V2 a) { return sqrt(pysq(a.x) + pysq(a.y)); }
static inline double length_sq(V2 a) { return pysq(a.x) + pysq(a.y); }
static inline V2 normalized(V2 a) {
  double l = length_(a);
  if (l < EPSILON) return a; /* common.py:216-222 (pure-Python flavour: DBL_EPSILON) */
  double il = 1.0 / l;
  return v2(a.x * il, a.y * il);
}
static inline double pmin(double a, double b) { return (b < a) ? b : a; } /* min(a, b) */
static inline double pmax(double a, double b) { return (b > a) ? b : a; } /* max(a, b) */
static inline double pclamp(double v, double lo, double hi) { return pmax(lo, pmin(v, hi)); }
static inline V2 vmin(V2 a, V2 b) { return v2(pmin(a.x, b.x), pmin(a.y, b.y)); }
static inline V2 vmax(V2 a, V2 b) { return v2(pmax(a.x, b.x), pmax(a.y, b.y)); }

/* PyMat22.angle setter (common.py:492-500) */
static inline M22 rot(double angle) {
  double c = cos(angle), s = sin(angle);
  M22 m; m.c1x = c; m.c2x = -s; m.c1y = s; m.c2y = c;
  return m;
}
static inline V2 mulmv(M22 m, V2 v) { /* common.py:405-412 */
  return v2(m.c1x * v.x + m.c2x * v.y, m.c1y * v.x + m.c2y * v.y);
}
static inline V2 mulmtv(M22 m, V2 v) { /* common.py:463-468: (col1.dot(v), col2.dot(v)) */
  return v2(m.c1x * v.x + m.c1y * v.y, m.c2x * v.x + m.c2y * v.y);
}
static inline M22 mulmtm(M22 a, M22 b) { /* common.py:478-482 */
  M22 r;
  r.c1x = a.c1x * b.c1x + a.c1y * b.c1y;
  r.c1y = a.c2x * b.c1x + a.c2y * b.c1y;
  r.c2x = a.c1x * b.c2x + a.c1y * b.c2y;
  r.c2y = a.c2x * b.c2x + a.c2y * b.c2y;
  return r;
}
/* PyTransform.__mul__ (common.py:677-685): (p.x + c1.x*o0) + c2.x*o1 */
static inline V2 xmul(XF t, V2 o) {
  return v2(t.p.x + t.R.c1x * o.x + t.R.c2x * o.y, t.p.y + t.R.c1y * o.x + t.R.c2y * o.y);
}
/* PyTransform.mul_t(vector) = R.mul_t(v - p) (common.py:666-669) */
static inline V2 xmult(XF t, V2 v) { return mulmtv(t.R, sub(v, t.p)); }
/* PyTransform.mul_t(transform) (common.py:670-673): quirk Q4 */
static inline XF xmult_xf(XF a, XF b) {
  XF r; r.p = sub(b.p, a.p); r.R = mulmtm(a.R, b.R); return r;
}
/* Transform(angle=a); position = c - R*lc  (body.py:906-908, contact.py:437-440) */
static inline XF xf_from(V2 c, double a, V2 lc) {
  XF t; t.R = rot(a); t.p = sub(c, mulmv(t.R, lc)); return t;
}
/* PyAABB.test_overlap (common.py:830-839) */
static inline int aabb_overlap(const AABB* a, const AABB* b) {
  V2 d1 = sub(b->lo, a->hi), d2 = sub(a->lo, b->hi);
  if (d1.x > 0.0 || d1.y > 0.0) return 0;
  if (d2.x > 0.0 || d2.y > 0.0) return 0;
  return 1;
}
/* PyAABB.contains (common.py:863-868) */
static inline int aabb_contains(const AABB* a, const AABB* b) {
  return (a->lo.x <= b->lo.x) && (a->lo.y <= b->lo.y) && (b->hi.x <= a->hi.x) && (b->hi.y <= a->hi.y);
}
static inline AABB aabb_union(AABB a, AABB b) { /* PyAABB.__add__ (common.py:857-861) */
  AABB r; r.lo = vmin(a.lo, b.lo); r.hi = vmax(a.hi, b.hi); return r;
}
/* Python list.index(min(xs)) / list.index(max(xs)): first index of the extreme */
static int argmin_first(const double* xs, int n) {
  double m = xs[0];
  for (int i = 1; i < n; ++i) if (xs[i] < m) m = xs[i];
  for (int i = 0; i < n; ++i) if (xs[i] == m) return i;
  return 0;
}
static int argmax_first(const double* xs, int n) {
  double m = xs[0];
  for (int i = 1; i < n; ++i) if (xs[i] > m) m = xs[i];
  for (int i = 0; i < n; ++i) if (xs[i] == m) return i;
  return 0;
}
static inline int pymod(int a, int n) { int r = a % n; return r < 0 ? r + n : r; }

/* ------------------------------------------------------------------ */
/* world state                                                         */
/* ------------------------------------------------------------------ */
typedef struct {
  int type;
  double radius;
  int n;
  V2 v[B2G_MAX_POLY_VERTS], nrm[B2G_MAX_POLY_VERTS];
  V2 centroid;
  V2 p;                  /* circle position */
  V2 e0, e1, e2, e3;     /* edge vertex0..vertex3 */
  int has0, has3;
} Shape;

typedef struct {
  int body;
  Shape sh;
  double friction, restitution, density;
  int category, mask, group, sensor;
  AABB fat;
} Fix;

typedef struct {
  int type, awake, allow_sleep, fixed_rot, active, bullet;
  int island_flag, island_index, island_id;
  XF xf;
  V2 lc, c0, c;
  double a0, a;
  V2 v; double w;
  V2 force; double torque;
  double mass, inv_mass, I, invI;
  double gscale, ldamp, adamp, sleep_time;
  double alpha0;  /* Sweep.alpha0 (common.py:870-923): the sweep's start within the step */
  int toi_flag;   /* in the TOI island being built */
} Body;

typedef struct { V2 lp; double ni, ti; int id[4]; } MPoint;
typedef struct { MPoint pts[2]; V2 ln, lp; int type, pc; } Manifold;

typedef struct {
  int fa, fb;
  double friction, restitution;
  Manifold m;
  int touching, enabled, filter, island_flag;
  int swapped; /* the registry swapped the reported pair (fa, fb) = (B, A) */
  double toi;                        /* cached time of impact (contact.py:64-81) */
  int toi_flag, toi_count, toi_island;
} Contact;

/* a joint: its record (b2g_layout.h B2G_JI_* / B2G_JD_*) stands for the joint object */
typedef struct {
  int32_t i[B2G_JOINT_NI];
  double d[B2G_JOINT_ND];
  int island_flag;
} Joint;

typedef struct {
  V2 gravity;
  double inv_dt0;
  int allow_sleep, warm_starting, new_fixture, keep_forces;
  int nb, nf, nc, ccap, nmove, mcap;
  Body* b;
  Fix* f;
  Contact* c;
  int* move;
  int* bfix_off; /* per-body fixture CSR (fixtures never change during stepping) */
  int* bfix;
  /* replay hint: unordered fixture-pair keys in reference creation order */
  int hint_n;
  int64_t* hint;
  int err;
  int index;              /* the world's index in its batch (event records) */
  /* contact listener: kinds recorded (1 << B2G_EV_*), the last step's records */
  int ev_mask, ev_n, ev_cap, post_key;
  int32_t* ev_i;
  double* ev_d;
  /* custom contact_filter: accept[a * nf + b] = filter(fixture a, fixture b); NULL: default */
  unsigned char* accept;
  int32_t* new_rows;      /* the contacts the last find_new_contacts created: fa, fb, swapped */
  int new_n, new_cap;
  /* a step in two parts (orc_step_begin / orc_step_end): its time step, and the list index
     after collide's removals of each contact's index during collide (-1: removed) */
  int split;
  double split_dt;
  int split_vi, split_pi;
  int* slot_map;
  int slot_n;
  /* continuous physics (World.solve_toi, world.py:594-821): on/off, listener records keyed in
     call order while it runs, and counts of time_of_impact calls and TOI events (sub-step
     island solves) in the last step */
  int continuous, toi_mode, toi_seq;
  long long toi_calls, toi_events;
  /* joints in World.joints order; per-body joint lists (Body.joints: the world's joints
     touching the body, in order, a joint listed twice when both its ends are the body) */
  int nj;
  Joint* j;
  int* bj_off;
  int* bj;
} World;

#define ORC_E_UNSUPPORTED 1
#define ORC_E_HINT 2
#define ORC_E_ALLOC 4
#define ORC_E_JOINT 16 /* a mouse joint's spring was too weak (the reference asserts, mouse.py:173) */

typedef struct { int nw; World* w; } Batch;

static void mp_id_unpack(int32_t packed, int id[4]) {
  uint32_t u = (uint32_t)packed;
  id[0] = (int)(u & 255u); id[1] = (int)((u >> 8) & 255u);
  id[2] = (int)((u >> 16) & 255u); id[3] = (int)((u >> 24) & 255u);
}
static int32_t mp_id_pack(const int id[4]) {
  uint32_t u = (uint32_t)(id[0] & 255) | ((uint32_t)(id[1] & 255) << 8) |
               ((uint32_t)(id[2] & 255) << 16) | ((uint32_t)(id[3] & 255) << 24);
  return (int32_t)u;
}
static inline void set_id(int id[4], int a, int b, int c, int d) { id[0] = a; id[1] = b; id[2] = c; id[3] = d; }

/* ------------------------------------------------------------------ */
/* shapes.py compute_aabb                                              */
/* ------------------------------------------------------------------ */
static AABB compute_aabb(const Shape* s, XF xf) {
  AABB r;
  if (s->type == B2G_SHAPE_CIRCLE) { /* shapes.py:178-189: p = xf.p + R*position */
    V2 p = add(xf.p, mulmv(xf.R, s->p));
    V2 rad = v2(s->radius, s->radius);
    r.lo = sub(p, rad); r.hi = add(p, rad);
  } else if (s->type == B2G_SHAPE_POLYGON) { /* shapes.py:316-334 */
    V2 lower = xmul(xf, s->v[0]);
    V2 upper = lower;
    for (int i = 0; i < s->n; ++i) {
      V2 v = xmul(xf, s->v[i]);
      lower = vmin(lower, v);
      upper = vmax(upper, v);
    }
    V2 rr = v2(s->radius, s->radius);
    r.lo = sub(lower, rr); r.hi = add(upper, rr);
  } else { /* edge, shapes.py:702-718 */
    V2 a = xmul(xf, s->e1), b = xmul(xf, s->e2);
    V2 rr = v2(s->radius, s->radius);
    r.lo = sub(vmin(a, b), rr); r.hi = add(vmax(a, b), rr);
  }
  return r;
}

/* ------------------------------------------------------------------ */
/* collision.py / collision_util.py                                    */
/* ------------------------------------------------------------------ */
typedef struct { V2 v; int id[4]; } ClipV; /* contact_util.py:97-118 */

static void collide_circles(Manifold* m, const Shape* A, XF xfa, const Shape* B, XF xfb) {
  /* collision.py:46-66 */
  m->pc = 0;
  V2 pa = xmul(xfa, A->p), pb = xmul(xfb, B->p);
  V2 d = sub(pb, pa);
  double dist_sqr = length_sq(d);
  double radius = A->radius + B->radius;
  if (dist_sqr > pysq(radius)) return;
  m->type = B2G_M_CIRCLES;
  m->lp = A->p;
  m->ln = v2(0.0, 0.0);
  m->pc = 1;
  m->pts[0].lp = B->p;
  set_id(m->pts[0].id, 0, 0, 0, 0);
}

static void collide_polygon_circle(Manifold* m, const Shape* P, XF xfa, const Shape* C, XF xfb) {
  /* collision.py:68-141 */
  m->pc = 0;
  V2 c = xmul(xfb, C->p);
  V2 cl = xmult(xfa, c);
  int normal_index = 0;
  double separation = -MAX_FLOAT;
  double radius = P->radius + C->radius;
  int n = P->n;
  for (int i = 0; i < n; ++i) {
    double s = dot(P->nrm[i], sub(cl, P->v[i]));
    if (s > radius) return;
    if (s > separation) { separation = s; normal_index = i; }
  }
  int i1 = normal_index, i2 = (i1 + 1) % n;
  V2 v1 = P->v[i1], v2_ = P->v[i2];
  if (separation < EPSILON) {
    m->pc = 1;
    m->type = B2G_M_FACE_A;
    m->ln = P->nrm[normal_index];
    m->lp = scale(0.5, add(v1, v2_));
    m->pts[0].lp = C->p;
    set_id(m->pts[0].id, 0, 0, 0, 0);
    return;
  }
  double u1 = dot(sub(cl, v1), sub(v2_, v1));
  double u2 = dot(sub(cl, v2_), sub(v1, v2_));
  if (u1 < 0.0) {
    V2 d = sub(cl, v1);
    if (dot(d, d) > pysq(radius)) return;
    m->ln = normalized(sub(cl, v1));
    m->lp = v1;
  } else if (u2 < 0.0) {
    V2 d = sub(cl, v2_);
    if (dot(d, d) > pysq(radius)) return;
    m->ln = normalized(sub(cl, v2_));
    m->lp = v2_;
  } else {
    V2 fc = scale(0.5, add(v1, v2_));
    separation = dot(sub(cl, fc), P->nrm[i1]);
    if (separation > radius) return;
    m->ln = P->nrm[i1];
    m->lp = fc;
  }
  m->pc = 1;
  m->type = B2G_M_FACE_A;
  m->pts[0].lp = C->p;
  set_id(m->pts[0].id, 0, 0, 0, 0);
}

/* collision_util.py:31-52 */
static double edge_separation(const Shape* p1, XF xf1, int edge1, const Shape* p2, XF xf2) {
  V2 n1w = mulmv(xf1.R, p1->nrm[edge1]);
  V2 n1 = mulmtv(xf2.R, n1w);
  double dots[B2G_MAX_POLY_VERTS];
  for (int i = 0; i < p2->n; ++i) dots[i] = dot(p2->v[i], n1);
  int index = argmin_first(dots, p2->n);
  V2 v1 = xmul(xf1, p1->v[edge1]);
  V2 v2_ = xmul(xf2, p2->v[index]);
  return dot(sub(v2_, v1), n1w);
}

/* collision_util.py:54-109 (hill climbing, quirk Q2/Q3) */
static double max_separation(const Shape* p1, XF xf1, const Shape* p2, XF xf2, int* edge_out) {
  int count1 = p1->n;
  V2 d = sub(xmul(xf2, p2->centroid), xmul(xf1, p1->centroid));
  V2 dl = mulmtv(xf1.R, d);
  double dots[B2G_MAX_POLY_VERTS];
  for (int i = 0; i < count1; ++i) dots[i] = dot(p1->nrm[i], dl);
  int edge = argmax_first(dots, count1);
  double s = edge_separation(p1, xf1, edge, p2, xf2);
  int prev_edge = pymod(edge - 1, count1);
  double s_prev = edge_separation(p1, xf1, prev_edge, p2, xf2);
  int next_edge = pymod(edge + 1, count1);
  double s_next = edge_separation(p1, xf1, next_edge, p2, xf2);
  int increment, best_edge;
  double best_sep;
  if (s_prev > s && s_prev > s_next) {
    increment = -1; best_edge = prev_edge; best_sep = s_prev;
  } else if (s_next > s) {
    increment = 1; best_edge = next_edge; best_sep = s_next;
  } else {
    *edge_out = edge;
    return s;
  }
  for (;;) {
    edge = increment == -1 ? pymod(best_edge - 1, count1) : pymod(best_edge + 1, count1);
    s = edge_separation(p1, xf1, edge, p2, xf2);
    if (s > best_sep) { best_edge = edge; best_sep = s; }
    else break;
  }
  *edge_out = best_edge;
  return best_sep;
}

/* collision_util.py:111-133 */
static void find_incident_edge(ClipV c[2], const Shape* p1, XF xf1, int edge1, const Shape* p2, XF xf2) {
  V2 n1 = mulmtv(xf2.R, mulmv(xf1.R, p1->nrm[edge1]));
  double dots[B2G_MAX_POLY_VERTS];
  for (int i = 0; i < p2->n; ++i) dots[i] = dot(n1, p2->nrm[i]);
  int i1 = argmin_first(dots, p2->n);
  int i2 = (i1 + 1) % p2->n;
  c[0].v = xmul(xf2, p2->v[i1]); set_id(c[0].id, edge1, i1, 1, 0);
  c[1].v = xmul(xf2, p2->v[i2]); set_id(c[1].id, edge1, i2, 1, 0);
}

/* collision_util.py:135-164 (Sutherland-Hodgman) */
static int clip_segment_to_line(ClipV out[3], const ClipV in[2], V2 normal, double offset, int via) {
  int n = 0;
  double d0 = dot(normal, in[0].v) - offset;
  double d1 = dot(normal, in[1].v) - offset;
  if (d0 <= 0.0) out[n++] = in[0];
  if (d1 <= 0.0) out[n++] = in[1];
  if (d0 * d1 < 0.0) {
    double interp = d0 / (d0 - d1);
    out[n].v = add(in[0].v, scale(interp, sub(in[1].v, in[0].v)));
    set_id(out[n].id, via, in[0].id[1], 0, 1);
    ++n;
  }
  return n;
}

static const double RELATIVE_TOL = 0.98, ABSOLUTE_TOL = 0.001; /* collision.py:36-37 */

static void collide_polygons(Manifold* m, const Shape* A, XF xfa, const Shape* B, XF xfb) {
  /* collision.py:144-241 */
  m->pc = 0;
  double total_radius = A->radius + B->radius;
  int edge_a, edge_b;
  double sa = max_separation(A, xfa, B, xfb, &edge_a);
  if (sa > total_radius) return;
  double sb = max_separation(B, xfb, A, xfa, &edge_b);
  if (sb > total_radius) return;
  const Shape *p1, *p2;
  XF xf1, xf2;
  int edge1, flip;
  if (sb > RELATIVE_TOL * sa + ABSOLUTE_TOL) {
    p1 = B; p2 = A; xf1 = xfb; xf2 = xfa; edge1 = edge_b; m->type = B2G_M_FACE_B; flip = 1;
  } else {
    p1 = A; p2 = B; xf1 = xfa; xf2 = xfb; edge1 = edge_a; m->type = B2G_M_FACE_A; flip = 0;
  }
  ClipV inc[2];
  find_incident_edge(inc, p1, xf1, edge1, p2, xf2);
  int iv1 = edge1, iv2 = (edge1 + 1) % p1->n;
  V2 v11 = p1->v[iv1], v12 = p1->v[iv2];
  V2 ltan = normalized(sub(v12, v11));
  V2 local_normal = cross_vs(ltan, 1.0);
  V2 planepoint = scale(0.5, add(v11, v12));
  V2 tangent = mulmv(xf1.R, ltan);
  V2 normal = cross_vs(tangent, 1.0);
  v11 = xmul(xf1, v11);
  v12 = xmul(xf1, v12);
  double front_offset = dot(normal, v11);
  double side1 = -dot(tangent, v11) + total_radius;
  double side2 = dot(tangent, v12) + total_radius;
  ClipV cp1[3], cp2[3];
  if (clip_segment_to_line(cp1, inc, neg(tangent), side1, iv1) < 2) return;
  if (clip_segment_to_line(cp2, cp1, tangent, side2, iv2) < 2) return;
  m->ln = local_normal;
  m->lp = planepoint;
  int pc = 0;
  for (int i = 0; i < 2; ++i) {
    double separation = dot(normal, cp2[i].v) - front_offset;
    if (separation <= total_radius) {
      MPoint* cp = &m->pts[pc];
      cp->lp = xmult(xf2, cp2[i].v);
      if (flip) set_id(cp->id, cp2[i].id[1], cp2[i].id[0], cp2[i].id[3], cp2[i].id[2]);
      else set_id(cp->id, cp2[i].id[0], cp2[i].id[1], cp2[i].id[2], cp2[i].id[3]);
      ++pc;
    }
  }
  m->pc = pc;
}

static void collide_edge_circle(Manifold* m, const Shape* E, XF xfa, const Shape* C, XF xfb) {
  /* collision.py:246-353 */
  m->pc = 0;
  V2 Q = xmult(xfa, xmul(xfb, C->p));
  V2 A = E->e1, B = E->e2;
  V2 e = sub(B, A);
  double u = dot(e, sub(B, Q));
  double v = dot(e, sub(Q, A));
  double radius = E->radius + C->radius;
  if (v <= 0.0) {
    V2 P = A, d = sub(Q, P);
    double dd = dot(d, d);
    if (dd > pysq(radius)) return;
    if (E->has0) {
      V2 A1 = E->e0, B1 = A;
      V2 e1 = sub(B1, A1);
      double u1 = dot(e1, sub(B1, Q));
      if (u1 > 0.0) return;
    }
    m->pc = 1; m->type = B2G_M_CIRCLES; m->ln = v2(0.0, 0.0); m->lp = P;
    set_id(m->pts[0].id, 0, 0, 0, 0);
    m->pts[0].lp = C->p;
    return;
  }
  if (u <= 0.0) {
    V2 P = B, d = sub(Q, P);
    double dd = dot(d, d);
    if (dd > pysq(radius)) return;
    if (E->has3) {
      V2 B2 = E->e3, A2 = B;
      V2 e2 = sub(B2, A2);
      double vv2 = dot(e2, sub(Q, A2));
      if (vv2 > 0.0) return;
    }
    m->pc = 1; m->type = B2G_M_CIRCLES; m->ln = v2(0.0, 0.0); m->lp = P;
    set_id(m->pts[0].id, 1, 0, 0, 0);
    m->pts[0].lp = C->p;
    return;
  }
  double den = dot(e, e);
  V2 P = scale(1.0 / den, add(scale(u, A), scale(v, B)));
  V2 d = sub(Q, P);
  double dd = dot(d, d);
  if (dd > pysq(radius)) return;
  V2 n = v2(-e.y, e.x);
  if (dot(n, sub(Q, A)) < 0.0) n = v2(-n.x, -n.y);
  n = normalized(n);
  m->pc = 1; m->type = B2G_M_FACE_A; m->ln = n; m->lp = A;
  set_id(m->pts[0].id, 0, 0, 1, 0);
  m->pts[0].lp = C->p;
}

/* EPCollider, collision.py:355-690 */
enum { EP_UNKNOWN = 0, EP_EDGE_A = 1, EP_EDGE_B = 2 };
typedef struct { int type, index; double sep; } EPAxis;
typedef struct { int n; V2 v[B2G_MAX_POLY_VERTS], nrm[B2G_MAX_POLY_VERTS]; V2 centroid; } EPProxy;
typedef struct {
  XF xf;
  V2 v0, v1, v2, v3, normal;
  int has0, has3;
  EPProxy pa, pb;
  double radius;
  V2 l11, l12, l21, l22;
} EPC;

static void ep_adjacency(EPC* c) { /* collision.py:541-604 */
  V2 v0 = c->v0, v1 = c->v1, vv2 = c->v2, v3 = c->v3;
  V2 cb = c->pb.centroid;
  if (c->has0) {
    V2 e0 = sub(v1, v0), e1 = sub(vv2, v1);
    V2 n0 = normalized(v2(e0.y, -e0.x)), n1 = normalized(v2(e1.y, -e1.x));
    int convex = cross(n0, n1) >= 0.0;
    int front0 = dot(n0, sub(cb, v0)) >= 0.0;
    int front1 = dot(n1, sub(cb, v1)) >= 0.0;
    if (convex) {
      if (front0 || front1) { c->l11 = n1; c->l12 = n0; }
      else { c->l11 = neg(n1); c->l12 = neg(n0); }
    } else {
      if (front0 && front1) { c->l11 = n0; c->l12 = n1; }
      else { c->l11 = neg(n0); c->l12 = neg(n1); }
    }
  } else {
    c->l11 = v2(0.0, 0.0); c->l12 = v2(0.0, 0.0);
  }
  if (c->has3) {
    V2 e1 = sub(vv2, v1), e2 = sub(v3, vv2);
    V2 n1 = normalized(v2(e1.y, -e1.x)), n2 = normalized(v2(e2.y, -e2.x));
    int convex = cross(n1, n2) >= 0.0;
    int front1 = dot(n1, sub(cb, v1)) >= 0.0;
    int front2 = dot(n2, sub(cb, vv2)) >= 0.0;
    if (convex) {
      if (front1 || front2) { c->l21 = n2; c->l22 = n1; }
      else { c->l21 = neg(n2); c->l22 = neg(n1); }
    } else {
      if (front1 && front2) { c->l21 = n1; c->l22 = n2; }
      else { c->l21 = neg(n1); c->l22 = neg(n2); }
    }
  } else {
    c->l21 = v2(0.0, 0.0); c->l22 = v2(0.0, 0.0);
  }
}

static inline int ep_valid(const EPC* c, V2 n) {
  int valid1 = (cross(n, c->l11) >= -ANGULAR_SLOP) && (cross(c->l12, n) >= -ANGULAR_SLOP);
  int valid2 = (cross(n, c->l21) >= -ANGULAR_SLOP) && (cross(c->l22, n) >= -ANGULAR_SLOP);
  return valid1 && valid2;
}

static EPAxis ep_edge_separation(const EPC* c) { /* collision.py:606-635 */
  EPAxis best; best.type = EP_UNKNOWN; best.index = -1; best.sep = -MAX_FLOAT;
  V2 normals[2] = {c->normal, neg(c->normal)};
  for (int i = 0; i < 2; ++i) {
    V2 n = normals[i];
    if (!ep_valid(c, n)) continue;
    EPAxis axis; axis.type = EP_EDGE_A; axis.index = i;
    double m = dot(n, sub(c->pb.v[0], c->v1));
    for (int k = 1; k < c->pb.n; ++k) {
      double d = dot(n, sub(c->pb.v[k], c->v1));
      if (d < m) m = d;
    }
    axis.sep = m;
    if (axis.sep > c->radius) return axis;
    if (axis.sep > best.sep) best = axis;
  }
  return best;
}

static EPAxis ep_polygon_separation(const EPC* c) { /* collision.py:637-666, quirk Q10 */
  EPAxis axis; axis.type = EP_UNKNOWN; axis.index = -1; axis.sep = -MAX_FLOAT;
  for (int i = 0; i < c->pb.n; ++i) {
    V2 n = neg(c->pb.nrm[i]);
    if (!ep_valid(c, n)) continue;
    double s1 = dot(n, sub(c->pb.v[i], c->v1));
    double s2 = dot(n, sub(c->pb.v[i], c->v2));
    double s = pmin(s1, s2);
    if (s > c->radius) { axis.type = EP_EDGE_B; axis.index = i; axis.sep = s; }
    if (s > axis.sep) { axis.type = EP_EDGE_B; axis.index = i; axis.sep = s; }
  }
  return axis;
}

static void ep_find_incident(ClipV out[2], const EPProxy* p1, int edge1, const EPProxy* p2) {
  /* collision.py:668-690 */
  V2 n1 = p1->nrm[edge1];
  double dots[B2G_MAX_POLY_VERTS];
  for (int i = 0; i < p2->n; ++i) dots[i] = dot(n1, p2->nrm[i]);
  int i1 = argmin_first(dots, p2->n);
  int i2 = (i1 + 1) % p2->n;
  out[0].v = p2->v[i1]; set_id(out[0].id, edge1, i1, 1, 0);
  out[1].v = p2->v[i2]; set_id(out[1].id, edge1, i2, 1, 0);
}

static void collide_edge_polygon(Manifold* m, const Shape* E, XF xfa, const Shape* P, XF xfb) {
  EPC c;
  c.xf = xmult_xf(xfa, xfb); /* collision.py:410 (Q4) */
  c.has0 = E->has0; c.has3 = E->has3;
  c.v0 = E->e0; c.v1 = E->e1; c.v2 = E->e2; c.v3 = E->e3;
  V2 e = sub(c.v2, c.v1);
  c.normal = normalized(v2(e.y, -e.x));
  c.pa.n = 2;
  c.pa.v[0] = c.v1; c.pa.v[1] = c.v2;
  c.pa.nrm[0] = c.normal; c.pa.nrm[1] = neg(c.normal);
  c.pa.centroid = scale(0.5, add(c.v1, c.v2));
  c.pb.centroid = xmul(c.xf, P->centroid);
  c.pb.n = P->n;
  for (int i = 0; i < P->n; ++i) {
    c.pb.v[i] = xmul(c.xf, P->v[i]);
    c.pb.nrm[i] = mulmv(c.xf.R, P->nrm[i]);
  }
  c.radius = 2.0 * POLYGON_RADIUS;

  /* EPCollider.collide, collision.py:442-535 */
  m->pc = 0;
  ep_adjacency(&c);
  EPAxis edge_axis = ep_edge_separation(&c);
  if (edge_axis.type == EP_UNKNOWN) return;
  if (edge_axis.sep > c.radius) return;
  EPAxis poly_axis = ep_polygon_separation(&c);
  if (poly_axis.type != EP_UNKNOWN && poly_axis.sep > c.radius) return;
  EPAxis primary;
  if (poly_axis.type == EP_UNKNOWN) primary = edge_axis;
  else if (poly_axis.sep > RELATIVE_TOL * edge_axis.sep + ABSOLUTE_TOL) primary = poly_axis;
  else primary = edge_axis;
  const EPProxy *p1, *p2;
  if (primary.type == EP_EDGE_A) { p1 = &c.pa; p2 = &c.pb; m->type = B2G_M_FACE_A; }
  else { p1 = &c.pb; p2 = &c.pa; m->type = B2G_M_FACE_B; }
  int edge1 = primary.index;
  ClipV inc[2];
  ep_find_incident(inc, p1, edge1, p2);
  int iv1 = edge1, iv2 = (edge1 + 1) % p1->n;
  V2 v11 = p1->v[iv1], v12 = p1->v[iv2];
  V2 tangent = normalized(sub(v12, v11));
  V2 normal = cross_vs(tangent, 1.0);
  V2 plane_point = scale(0.5, add(v11, v12));
  double front_offset = dot(normal, v11);
  double side1 = -dot(tangent, v11) + c.radius;
  double side2 = dot(tangent, v12) + c.radius;
  ClipV cp1[3], cp2[3];
  if (clip_segment_to_line(cp1, inc, neg(tangent), side1, iv1) < 2) return;
  int n2 = clip_segment_to_line(cp2, cp1, tangent, side2, iv2);
  if (n2 < 2) return;
  if (primary.type == EP_EDGE_A) { m->ln = normal; m->lp = plane_point; }
  else { m->ln = mulmtv(c.xf.R, normal); m->lp = xmult(c.xf, plane_point); }
  int pc = 0;
  for (int i = 0; i < n2; ++i) {
    double separation = dot(normal, cp2[i].v) - front_offset;
    if (separation <= c.radius) {
      MPoint* mp = &m->pts[pc];
      if (primary.type == EP_EDGE_A) {
        mp->lp = xmult(c.xf, cp2[i].v);
        set_id(mp->id, cp2[i].id[0], cp2[i].id[1], cp2[i].id[2], cp2[i].id[3]);
      } else {
        mp->lp = cp2[i].v;
        /* set_id(type_b, type_a, index_b, index_a), quirk Q9 */
        set_id(mp->id, cp2[i].id[3], cp2[i].id[2], cp2[i].id[1], cp2[i].id[0]);
      }
      ++pc;
    }
  }
  m->pc = pc;
}

/* Contact.evaluate dispatch (contact.py:343-373); fixtures are already
   oriented as the contact registry orders them (contact.py:305-326). */
static int evaluate(Manifold* m, const Shape* A, XF xfa, const Shape* B, XF xfb) {
  int ta = A->type, tb = B->type;
  if (ta == B2G_SHAPE_CIRCLE && tb == B2G_SHAPE_CIRCLE) collide_circles(m, A, xfa, B, xfb);
  else if (ta == B2G_SHAPE_POLYGON && tb == B2G_SHAPE_POLYGON) collide_polygons(m, A, xfa, B, xfb);
  else if (ta == B2G_SHAPE_POLYGON && tb == B2G_SHAPE_CIRCLE) collide_polygon_circle(m, A, xfa, B, xfb);
  else if (ta == B2G_SHAPE_EDGE && tb == B2G_SHAPE_CIRCLE) collide_edge_circle(m, A, xfa, B, xfb);
  else if (ta == B2G_SHAPE_EDGE && tb == B2G_SHAPE_POLYGON) collide_edge_polygon(m, A, xfa, B, xfb);
  else return -1;
  return 0;
}

/* ------------------------------------------------------------------ */
/* body.py helpers                                                     */
/* ------------------------------------------------------------------ */
static void set_awake(Body* b, int awake) { /* body.py:800-812 */
  if (awake) {
    if (!b->awake) b->sleep_time = 0.0;
  } else {
    b->sleep_time = 0.0;
    b->v = v2(0.0, 0.0);
    b->w = 0.0;
    b->force = v2(0.0, 0.0);
    b->torque = 0.0;
  }
  b->awake = awake;
}

static inline int other_body(const World* w, const Contact* c, int body) { /* contact.py:92-99 */
  int ba = w->f[c->fa].body;
  return body == ba ? w->f[c->fb].body : ba;
}

/* ------------------------------------------------------------------ */
/* broad-phase + ContactManager.find_new_contacts                      */
/* ------------------------------------------------------------------ */
typedef struct { int64_t* keys; int cap; } HSet;
static int hs_init(HSet* h, int n) {
  int cap = 16;
  while (cap < 2 * n + 16) cap <<= 1;
  h->keys = (int64_t*)malloc(sizeof(int64_t) * (size_t)cap);
  if (!h->keys) return -1;
  for (int i = 0; i < cap; ++i) h->keys[i] = -1;
  h->cap = cap;
  return 0;
}
static inline uint64_t hs_hash(int64_t k) {
  uint64_t x = (uint64_t)k * 0x9E3779B97F4A7C15ull;
  return x ^ (x >> 29);
}
static int hs_has(const HSet* h, int64_t k) {
  uint64_t i = hs_hash(k) & (uint64_t)(h->cap - 1);
  while (h->keys[i] != -1) {
    if (h->keys[i] == k) return 1;
    i = (i + 1) & (uint64_t)(h->cap - 1);
  }
  return 0;
}
static int hs_add(HSet* h, int64_t k) { /* returns 1 if inserted; caller sizes the set */
  uint64_t i = hs_hash(k) & (uint64_t)(h->cap - 1);
  while (h->keys[i] != -1) {
    if (h->keys[i] == k) return 0;
    i = (i + 1) & (uint64_t)(h->cap - 1);
  }
  h->keys[i] = k;
  return 1;
}
static inline int64_t okey(int a, int b) { return ((int64_t)a << 32) | (int64_t)(uint32_t)b; }
static inline int64_t ukey(int a, int b) { return a < b ? okey(a, b) : okey(b, a); }

/* BroadPhase.update_pairs (broadphase.py:83-109) restated over all proxies:
   the tree query returns exactly the proxies whose fat AABB overlaps the
   query's (dynamictree.py:190-207; pinned by dynamictree_test.py:89-98). */
static int update_pairs(const AABB* fat, int nf, const int* move, int nmove, int** out_pairs, int* out_n) {
  int cap = 64, n = 0;
  int* pairs = (int*)malloc(sizeof(int) * 2 * (size_t)cap);
  HSet set;
  if (!pairs || hs_init(&set, nmove * 8 + 64)) { free(pairs); return -1; }
  for (int k = 0; k < nmove; ++k) {
    int P = move[k];
    for (int R = 0; R < nf; ++R) {
      if (R == P) continue;
      if (!aabb_overlap(&fat[R], &fat[P])) continue;
      if (hs_has(&set, okey(R, P))) continue;
      if (2 * (n + 1) > set.cap) { /* grow the set */
        HSet bigger;
        if (hs_init(&bigger, set.cap)) { free(set.keys); free(pairs); return -1; }
        for (int i = 0; i < set.cap; ++i) if (set.keys[i] != -1) hs_add(&bigger, set.keys[i]);
        free(set.keys);
        set = bigger;
      }
      if (!hs_add(&set, okey(P, R))) continue;
      if (n == cap) {
        cap *= 2;
        int* np = (int*)realloc(pairs, sizeof(int) * 2 * (size_t)cap);
        if (!np) { free(set.keys); free(pairs); return -1; }
        pairs = np;
      }
      pairs[2 * n] = P;
      pairs[2 * n + 1] = R;
      ++n;
    }
  }
  free(set.keys);
  *out_pairs = pairs;
  *out_n = n;
  return 0;
}

static int contact_reserve(World* w, int n) {
  if (n <= w->ccap) return 0;
  int cap = w->ccap ? w->ccap : 16;
  while (cap < n) cap *= 2;
  Contact* nc = (Contact*)realloc(w->c, sizeof(Contact) * (size_t)cap);
  if (!nc) return -1;
  w->c = nc;
  w->ccap = cap;
  return 0;
}

static int64_t* g_sort_rank_keys; /* used by the hint sort */
static int g_sort_rank_n;
static int hint_rank(int64_t key) {
  for (int i = 0; i < g_sort_rank_n; ++i) if (g_sort_rank_keys[i] == key) return i;
  return g_sort_rank_n;
}

/* One listener call as an event record (b2g_layout.h): the contact's pair, flags and
   manifold at the call.  Returns the record's index, or -1. */
static void ev_manifold(const Manifold* m, double* o, int32_t* oi) {
  o[0] = m->ln.x; o[1] = m->ln.y; o[2] = m->lp.x; o[3] = m->lp.y;
  o[4] = m->pts[0].lp.x; o[5] = m->pts[0].lp.y; o[6] = m->pts[0].ni; o[7] = m->pts[0].ti;
  o[8] = m->pts[1].lp.x; o[9] = m->pts[1].lp.y; o[10] = m->pts[1].ni; o[11] = m->pts[1].ti;
  oi[0] = m->type;
  oi[1] = m->pc;
  oi[2] = mp_id_pack(m->pts[0].id);
  oi[3] = mp_id_pack(m->pts[1].id);
}

static int ev_push(World* w, int kind, int phase, int key, const Contact* c) {
  if (!(w->ev_mask & (1 << kind))) return -1;
  if (w->toi_mode) { /* inside World.solve_toi: records come after the step's others, in call order */
    phase = 3;
    key = w->toi_seq++;
  }
  if (w->ev_n == w->ev_cap) {
    int cap = w->ev_cap ? 2 * w->ev_cap : 64;
    int32_t* ni = (int32_t*)realloc(w->ev_i, sizeof(int32_t) * B2G_EV_NI * (size_t)cap);
    if (ni) w->ev_i = ni;
    double* nd = (double*)realloc(w->ev_d, sizeof(double) * B2G_EV_ND * (size_t)cap);
    if (nd) w->ev_d = nd;
    if (!ni || !nd) { w->err |= ORC_E_ALLOC; return -1; }
    w->ev_cap = cap;
  }
  int k = w->ev_n++;
  int32_t* ei = w->ev_i + (size_t)k * B2G_EV_NI;
  double* ed = w->ev_d + (size_t)k * B2G_EV_ND;
  memset(ei, 0, sizeof(int32_t) * B2G_EV_NI);
  memset(ed, 0, sizeof(double) * B2G_EV_ND);
  ei[B2G_EI_WORLD] = w->index;
  ei[B2G_EI_KIND] = kind;
  ei[B2G_EI_PHASE] = phase;
  ei[B2G_EI_KEY] = key;
  ei[B2G_EI_FA] = c->fa;
  ei[B2G_EI_FB] = c->fb;
  ei[B2G_EI_FLAGS] = (c->touching ? B2G_CF_TOUCHING : 0) | (c->enabled ? B2G_CF_ENABLED : 0);
  ev_manifold(&c->m, ed, ei + B2G_EI_MTYPE);
  return k;
}

/* Body._should_collide's joint test (body.py:920-926): a joint between the two bodies that does
   not let them collide */
static int joint_blocks(const World* w, int ba, int bb) {
  for (int k = w->bj_off[bb]; k < w->bj_off[bb + 1]; ++k) {
    const Joint* J = &w->j[w->bj[k]];
    int a = J->i[B2G_JI_BODY_A], b = J->i[B2G_JI_BODY_B];
    if ((a == bb && b == ba) || (b == bb && a == ba))
      if (!(J->i[B2G_JI_FLAGS] & B2G_JF_COLLIDE_CONNECTED)) return 1;
  }
  return 0;
}

/* contact_filter(fixture a, fixture b): the custom filter's verdict, or World.default_contact_filter
   (world.py:958-970) */
static int filter_ok(const World* w, int a, int b) {
  if (w->accept) return w->accept[(size_t)a * (size_t)w->nf + (size_t)b] != 0;
  const Fix *A = &w->f[a], *B = &w->f[b];
  if (A->group == B->group && A->group != 0) return A->group > 0;
  return ((A->mask & B->category) != 0) && ((A->category & B->mask) != 0);
}

static void find_new_contacts(World* w) { /* contact_manager.py:102-158 */
  int* pairs = NULL;
  int np = 0;
  AABB* fat = (AABB*)malloc(sizeof(AABB) * (size_t)(w->nf > 0 ? w->nf : 1));
  if (!fat) { w->err |= ORC_E_ALLOC; return; }
  for (int i = 0; i < w->nf; ++i) fat[i] = w->f[i].fat;
  if (update_pairs(fat, w->nf, w->move, w->nmove, &pairs, &np)) { free(fat); w->err |= ORC_E_ALLOC; return; }
  free(fat);
  w->nmove = 0;
  HSet exist;
  if (hs_init(&exist, w->nc)) { free(pairs); w->err |= ORC_E_ALLOC; return; }
  for (int i = 0; i < w->nc; ++i) hs_add(&exist, ukey(w->c[i].fa, w->c[i].fb));
  int first_new = w->nc;
  for (int k = 0; k < np; ++k) {
    int fa = pairs[2 * k], fb = pairs[2 * k + 1];
    int ba = w->f[fa].body, bb = w->f[fb].body;
    if (ba == bb) continue;
    if (hs_has(&exist, ukey(fa, fb))) continue;
    /* body_b._should_collide(body_a) (body.py:910-928) */
    if (w->b[bb].type != B2G_DYNAMIC && w->b[ba].type != B2G_DYNAMIC) continue;
    if (joint_blocks(w, ba, bb)) continue;
    /* the contact filter (contact_manager.py:144-146) */
    const Fix *A = &w->f[fa], *B = &w->f[fb];
    if (!filter_ok(w, fa, fb)) continue;
    /* Contact._create + registry (contact.py:294-326) */
    int ta = A->sh.type, tb = B->sh.type, swap;
    if (ta == B2G_SHAPE_LOOP || tb == B2G_SHAPE_LOOP) { w->err |= ORC_E_UNSUPPORTED; continue; }
    if ((ta == 0 && tb == 0) || (ta == 1 && tb == 1) || (ta == 1 && tb == 0) || (ta == 3 && tb == 0) ||
        (ta == 3 && tb == 1)) swap = 0;
    else if ((ta == 0 && tb == 1) || (ta == 0 && tb == 3) || (ta == 1 && tb == 3)) swap = 1;
    else continue; /* no collider registered (edge/edge) */
    if (contact_reserve(w, w->nc + 1)) { w->err |= ORC_E_ALLOC; break; }
    Contact* c = &w->c[w->nc++];
    memset(c, 0, sizeof(*c));
    c->fa = swap ? fb : fa;
    c->fb = swap ? fa : fb;
    c->swapped = swap;
    /* mix_friction / mix_restitution (contact_util.py:36-48) */
    c->friction = sqrt(w->f[c->fa].friction * w->f[c->fb].friction);
    c->restitution = pmax(w->f[c->fa].restitution, w->f[c->fb].restitution);
    c->m.type = B2G_M_CIRCLES;
  }
  free(exist.keys);
  free(pairs);
  if (w->hint_n > 0 && w->nc > first_new) {
    /* replay: order this call's new contacts by their rank in the hint */
    int nn = w->nc - first_new;
    g_sort_rank_keys = w->hint;
    g_sort_rank_n = w->hint_n;
    int* rank = (int*)malloc(sizeof(int) * (size_t)nn);
    Contact* tmp = (Contact*)malloc(sizeof(Contact) * (size_t)nn);
    if (!rank || !tmp) { free(rank); free(tmp); w->err |= ORC_E_ALLOC; return; }
    for (int i = 0; i < nn; ++i) {
      rank[i] = hint_rank(ukey(w->c[first_new + i].fa, w->c[first_new + i].fb));
      if (rank[i] == w->hint_n) w->err |= ORC_E_HINT;
      tmp[i] = w->c[first_new + i];
    }
    for (int i = 1; i < nn; ++i) { /* stable insertion sort by rank */
      Contact ct = tmp[i];
      int r = rank[i], j = i - 1;
      while (j >= 0 && rank[j] > r) { tmp[j + 1] = tmp[j]; rank[j + 1] = rank[j]; --j; }
      tmp[j + 1] = ct; rank[j + 1] = r;
    }
    memcpy(&w->c[first_new], tmp, sizeof(Contact) * (size_t)nn);
    free(rank);
    free(tmp);
  }
  w->new_n = 0;
  int nn = w->nc - first_new;
  if (nn > w->new_cap) {
    int32_t* r = (int32_t*)realloc(w->new_rows, sizeof(int32_t) * 3 * (size_t)nn);
    if (!r) { w->err |= ORC_E_ALLOC; return; }
    w->new_rows = r;
    w->new_cap = nn;
  }
  for (int i = 0; i < nn; ++i) {
    const Contact* c = &w->c[first_new + i];
    w->new_rows[3 * i] = c->fa;
    w->new_rows[3 * i + 1] = c->fb;
    w->new_rows[3 * i + 2] = c->swapped ? (int32_t)B2G_CF_SWAPPED : 0;
  }
  w->new_n = nn;
}

/* ------------------------------------------------------------------ */
/* ContactManager.collide + Contact.update                             */
/* ------------------------------------------------------------------ */
static void contact_update(World* w, Contact* c, int slot) { /* contact.py:210-268 */
  Manifold old = c->m;
  c->enabled = 1;
  int was_touching = c->touching;
  Fix *FA = &w->f[c->fa], *FB = &w->f[c->fb];
  if (FA->sensor || FB->sensor) { w->err |= ORC_E_UNSUPPORTED; return; }
  Body *A = &w->b[FA->body], *B = &w->b[FB->body];
  if (evaluate(&c->m, &FA->sh, A->xf, &FB->sh, B->xf)) { w->err |= ORC_E_UNSUPPORTED; return; }
  int touching = c->m.pc > 0;
  for (int j = 0; j < c->m.pc; ++j) {
    MPoint* mp2 = &c->m.pts[j];
    int found = -1;
    /* dict(...) keeps the last duplicate key */
    for (int i = 0; i < old.pc; ++i)
      if (memcmp(old.pts[i].id, mp2->id, sizeof(mp2->id)) == 0) found = i;
    if (found >= 0) { mp2->ni = old.pts[found].ni; mp2->ti = old.pts[found].ti; }
    else { mp2->ni = 0.0; mp2->ti = 0.0; }
  }
  if (touching != was_touching) { set_awake(A, 1); set_awake(B, 1); }
  c->touching = touching;
  /* the listener calls (contact.py:262-268), keyed by the contact's place in the list */
  if (!was_touching && touching) ev_push(w, B2G_EV_BEGIN, 0, 4 * slot, c);
  if (was_touching && !touching) ev_push(w, B2G_EV_END, 0, 4 * slot + 1, c);
  if (touching) {
    int k = ev_push(w, B2G_EV_PRE_SOLVE, 0, 4 * slot + 2, c);
    if (k >= 0) ev_manifold(&old, w->ev_d + (size_t)k * B2G_EV_ND + B2G_ED_OLD,
                            w->ev_i + (size_t)k * B2G_EV_NI + B2G_EI_OLD_MTYPE);
  }
}

static void collide(World* w) { /* contact_manager.py:174-223 */
  char* remove = (char*)calloc((size_t)(w->nc > 0 ? w->nc : 1), 1);
  if (!remove) { w->err |= ORC_E_ALLOC; return; }
  int nremove = 0;
  for (int i = 0; i < w->nc; ++i) {
    Contact* c = &w->c[i];
    Fix *FA = &w->f[c->fa], *FB = &w->f[c->fb];
    Body *A = &w->b[FA->body], *B = &w->b[FB->body];
    int active_a = A->awake && A->type != B2G_STATIC;
    int active_b = B->awake && B->type != B2G_STATIC;
    if (!active_a && !active_b) continue;
    if (c->filter) {
      int ok = (A->type == B2G_DYNAMIC || B->type == B2G_DYNAMIC);
      if (ok) ok = !joint_blocks(w, FA->body, FB->body);
      if (ok) ok = filter_ok(w, c->fa, c->fb);
      if (!ok) { remove[i] = 1; ++nremove; continue; }
      c->filter = 0;
    }
    if (!aabb_overlap(&FA->fat, &FB->fat)) { remove[i] = 1; ++nremove; continue; }
    contact_update(w, c, i);
  }
  for (int i = 0; i < w->nc; ++i) /* ContactManager.destroy: end_contact (contact_manager.py:160-172) */
    if (remove[i] && w->c[i].touching) ev_push(w, B2G_EV_END, 1, i, &w->c[i]);
  free(w->slot_map);
  w->slot_map = (int*)malloc(sizeof(int) * (size_t)(w->nc > 0 ? w->nc : 1));
  w->slot_n = w->slot_map ? w->nc : 0;
  if (!w->slot_map) w->err |= ORC_E_ALLOC;
  { /* ContactManager.destroy keeps the order of the others */
    int k = 0;
    for (int i = 0; i < w->nc; ++i) {
      if (w->slot_map) w->slot_map[i] = remove[i] ? -1 : k;
      if (!remove[i]) w->c[k++] = w->c[i];
    }
    w->nc = k;
  }
  free(remove);
}

/* ------------------------------------------------------------------ */
/* ContactSolver (contact.py:375-864) and Island.solve (island.py)     */
/* ------------------------------------------------------------------ */
typedef struct { V2 ra, rb; double ni, ti, nm, tm, bias; } VCP;
typedef struct {
  int npts;
  VCP p[2];
  V2 normal;
  M22 K, nmass;
  int ia, ib;
  double ma, mb, Ia, Ib, friction, restitution;
  int contact;
} VC;
typedef struct {
  int npts;
  V2 lpts[2], ln, lp;
  int type, ia, ib;
  double ra, rb, ma, mb, Ia, Ib;
  V2 lca, lcb;
} PC;
typedef struct { V2 c; double a; V2 v; double w; } PV;

typedef struct { double dt, inv_dt, dt_ratio; int vel_iters, pos_iters, warm_starting; } Step;

/* WorldManifold (contact_util.py:195-245) */
static V2 world_manifold(const Manifold* m, XF xfa, double ra, XF xfb, double rb, V2 pts[2]) {
  V2 normal = v2(0.0, 0.0);
  if (m->pc == 0) return normal;
  if (m->type == B2G_M_CIRCLES) {
    normal = v2(1.0, 0.0);
    V2 pa = xmul(xfa, m->lp);
    V2 pb = xmul(xfb, m->pts[0].lp);
    V2 d = sub(pa, pb);
    if (dot(d, d) > EPSILON_SQR) normal = normalized(sub(pb, pa));
    V2 ca = add(pa, scale(ra, normal));
    V2 cb = sub(pb, scale(rb, normal));
    pts[0] = scale(0.5, add(ca, cb));
  } else if (m->type == B2G_M_FACE_A) {
    normal = mulmv(xfa.R, m->ln);
    V2 plane = xmul(xfa, m->lp);
    for (int i = 0; i < m->pc; ++i) {
      V2 clip = xmul(xfb, m->pts[i].lp);
      V2 ca = add(clip, scale(ra - dot(sub(clip, plane), normal), normal));
      V2 cb = sub(clip, scale(rb, normal));
      pts[i] = scale(0.5, add(ca, cb));
    }
  } else {
    normal = mulmv(xfb.R, m->ln);
    V2 plane = xmul(xfb, m->lp);
    for (int i = 0; i < m->pc; ++i) {
      V2 clip = xmul(xfa, m->pts[i].lp);
      V2 cb = add(clip, scale(rb - dot(sub(clip, plane), normal), normal));
      V2 ca = sub(clip, scale(ra, normal));
      pts[i] = scale(0.5, add(ca, cb));
    }
    normal = neg(normal);
  }
  return normal;
}

/* ContactPositionConstraint.solver_manifold (contact_util.py:323-360) */
static void solver_manifold(const PC* pc, XF xfa, XF xfb, int j, V2* normal, V2* point, double* sep) {
  if (pc->type == B2G_M_CIRCLES) {
    V2 pa = xmul(xfa, pc->lp), pb = xmul(xfb, pc->lpts[0]);
    *normal = normalized(sub(pb, pa));
    *point = scale(0.5, add(pa, pb));
    *sep = dot(sub(pb, pa), *normal) - pc->ra - pc->rb;
  } else if (pc->type == B2G_M_FACE_A) {
    *normal = mulmv(xfa.R, pc->ln);
    V2 plane = xmul(xfa, pc->lp);
    V2 clip = xmul(xfb, pc->lpts[j]);
    *sep = dot(sub(clip, plane), *normal) - pc->ra - pc->rb;
    *point = clip;
  } else {
    *normal = mulmv(xfb.R, pc->ln);
    V2 plane = xmul(xfb, pc->lp);
    V2 clip = xmul(xfa, pc->lpts[j]);
    *sep = dot(sub(clip, plane), *normal) - pc->ra - pc->rb;
    *point = clip;
    *normal = neg(*normal);
  }
}

static inline void apply_block(VC* vc, V2 x, V2 a, V2 normal, V2* va, double* wa, V2* vb, double* wb) {
  V2 d = sub(x, a);
  V2 p1 = scale(d.x, normal), p2 = scale(d.y, normal);
  *va = sub(*va, scale(vc->ma, add(p1, p2)));
  *wa -= vc->Ia * (cross(vc->p[0].ra, p1) + cross(vc->p[1].ra, p2));
  *vb = add(*vb, scale(vc->mb, add(p1, p2)));
  *wb += vc->Ib * (cross(vc->p[0].rb, p1) + cross(vc->p[1].rb, p2));
  vc->p[0].ni = x.x;
  vc->p[1].ni = x.y;
}

static void solve_velocity(VC* vcs, int nvc, PV* pv) { /* contact.py:527-799 */
  for (int k = 0; k < nvc; ++k) {
    VC* vc = &vcs[k];
    V2 va = pv[vc->ia].v, vb = pv[vc->ib].v;
    double wa = pv[vc->ia].w, wb = pv[vc->ib].w;
    V2 normal = vc->normal;
    V2 tangent = cross_vs(normal, 1.0);
    double friction = vc->friction;
    for (int j = 0; j < vc->npts; ++j) {
      VCP* p = &vc->p[j];
      V2 dv = sub(sub(add(vb, cross_sv(wb, p->rb)), va), cross_sv(wa, p->ra));
      double vt = dot(dv, tangent);
      double lambda = p->tm * (-vt);
      double maxf = friction * p->ni;
      double newi = pclamp(p->ti + lambda, -maxf, maxf);
      lambda = newi - p->ti;
      V2 P = scale(lambda, tangent);
      va = sub(va, scale(vc->ma, P));
      wa -= vc->Ia * cross(p->ra, P);
      vb = add(vb, scale(vc->mb, P));
      wb += vc->Ib * cross(p->rb, P);
      p->ti = newi;
    }
    if (vc->npts == 1) {
      VCP* p = &vc->p[0];
      V2 dv = sub(sub(add(vb, cross_sv(wb, p->rb)), va), cross_sv(wa, p->ra));
      double vn = dot(dv, normal);
      double lambda = -p->nm * (vn - p->bias);
      double newi = pmax(p->ni + lambda, 0.0);
      lambda = newi - p->ni;
      V2 P = scale(lambda, normal);
      va = sub(va, scale(vc->ma, P));
      wa -= vc->Ia * cross(p->ra, P);
      vb = add(vb, scale(vc->mb, P));
      wb += vc->Ib * cross(p->rb, P);
      p->ni = newi;
    } else {
      VCP *cp1 = &vc->p[0], *cp2 = &vc->p[1];
      V2 a = v2(cp1->ni, cp2->ni);
      V2 dv1 = sub(sub(add(vb, cross_sv(wb, cp1->rb)), va), cross_sv(wa, cp1->ra));
      V2 dv2 = sub(sub(add(vb, cross_sv(wb, cp2->rb)), va), cross_sv(wa, cp2->ra));
      double vn1 = dot(dv1, normal);
      double vn2 = dot(dv2, normal);
      V2 b = v2(vn1 - cp1->bias, vn2 - cp2->bias);
      b = sub(b, mulmv(vc->K, a));
      for (;;) {
        V2 x = neg(mulmv(vc->nmass, b)); /* case 1 */
        if (x.x >= 0.0 && x.y >= 0.0) { apply_block(vc, x, a, normal, &va, &wa, &vb, &wb); break; }
        x = v2(-cp1->nm * b.x, 0.0); /* case 2 */
        vn2 = vc->K.c1y * x.x + b.y;
        if (x.x >= 0.0 && vn2 >= 0.0) { apply_block(vc, x, a, normal, &va, &wa, &vb, &wb); break; }
        x = v2(0.0, -cp2->nm * b.y); /* case 3 */
        vn1 = vc->K.c2x * x.y + b.x;
        if (x.y >= 0.0 && vn1 >= 0.0) { apply_block(vc, x, a, normal, &va, &wa, &vb, &wb); break; }
        x = v2(0.0, 0.0); /* case 4 */
        vn1 = b.x;
        vn2 = b.y;
        if (vn1 >= 0.0 && vn2 >= 0.0) { apply_block(vc, x, a, normal, &va, &wa, &vb, &wb); break; }
        break;
      }
    }
    pv[vc->ia].v = va; pv[vc->ia].w = wa;
    pv[vc->ib].v = vb; pv[vc->ib].w = wb;
  }
}

static int solve_position(PC* pcs, int npc, PV* pv) { /* contact.py:809-864 */
  double min_sep = 0.0;
  for (int k = 0; k < npc; ++k) {
    PC* pc = &pcs[k];
    V2 ca = pv[pc->ia].c, cb = pv[pc->ib].c;
    double aa = pv[pc->ia].a, ab = pv[pc->ib].a;
    for (int j = 0; j < pc->npts; ++j) {
      XF xfa = xf_from(ca, aa, pc->lca);
      XF xfb = xf_from(cb, ab, pc->lcb);
      V2 normal, point;
      double sep;
      solver_manifold(pc, xfa, xfb, j, &normal, &point, &sep);
      V2 ra = sub(point, ca), rb = sub(point, cb);
      min_sep = pmin(min_sep, sep);
      double C = pclamp(BAUMGARTE * (sep + LINEAR_SLOP), -MAX_LINEAR_CORRECTION, 0.0);
      double rna = pysq(cross(ra, normal));
      double rnb = pysq(cross(rb, normal));
      double K = pc->ma + pc->mb + (pc->Ia * rna) + (pc->Ib * rnb);
      double impulse = K > 0.0 ? -C / K : 0.0;
      V2 P = scale(impulse, normal);
      ca = sub(ca, scale(pc->ma, P));
      aa -= pc->Ia * cross(ra, P);
      cb = add(cb, scale(pc->mb, P));
      ab += pc->Ib * cross(rb, P);
    }
    pv[pc->ia].c = ca; pv[pc->ia].a = aa;
    pv[pc->ib].c = cb; pv[pc->ib].a = ab;
  }
  return min_sep >= -3.0 * LINEAR_SLOP;
}

typedef struct {
  int* bodies; int nbodies;
  int* contacts; int ncontacts;
  int* joints; int njoints;
  PV* pv; VC* vc; PC* pc;
} IslandBuf;

/* ContactSolver.__init__ (contact.py:376-412) over the island's contacts */
static void cs_build(World* w, IslandBuf* is, const Step* st) {
  int nvc = is->ncontacts;
  for (int k = 0; k < nvc; ++k) {
    Contact* c = &w->c[is->contacts[k]];
    Fix *FA = &w->f[c->fa], *FB = &w->f[c->fb];
    Body *A = &w->b[FA->body], *B = &w->b[FB->body];
    VC* vc = &is->vc[k];
    PC* pc = &is->pc[k];
    memset(vc, 0, sizeof(*vc));
    vc->contact = is->contacts[k];
    vc->friction = c->friction;
    vc->restitution = c->restitution;
    vc->ia = A->island_index; vc->ib = B->island_index;
    vc->ma = A->inv_mass; vc->mb = B->inv_mass;
    vc->Ia = A->invI; vc->Ib = B->invI;
    vc->K.c1x = 1.0; vc->K.c2y = 1.0;
    vc->nmass.c1x = 1.0; vc->nmass.c2y = 1.0;
    vc->npts = c->m.pc;
    for (int j = 0; j < c->m.pc; ++j) { /* VelocityConstraintPoint (contact_util.py:253-277) */
      if (st->warm_starting) {
        vc->p[j].ni = st->dt_ratio * c->m.pts[j].ni;
        vc->p[j].ti = st->dt_ratio * c->m.pts[j].ti;
      } else {
        vc->p[j].ni = c->m.pts[j].ni;
        vc->p[j].ti = c->m.pts[j].ti;
      }
    }
    pc->ia = vc->ia; pc->ib = vc->ib;
    pc->ma = vc->ma; pc->mb = vc->mb;
    pc->Ia = vc->Ia; pc->Ib = vc->Ib;
    pc->lca = A->lc; pc->lcb = B->lc;
    pc->ra = FA->sh.radius; pc->rb = FB->sh.radius;
    pc->ln = c->m.ln; pc->lp = c->m.lp;
    pc->type = c->m.type;
    pc->npts = c->m.pc;
    for (int j = 0; j < c->m.pc; ++j) pc->lpts[j] = c->m.pts[j].lp;
  }
}

/* ContactSolver.initialize_velocity_constraints (contact.py:414-496) at the island's positions */
static void cs_init_velocity(World* w, IslandBuf* is) {
  PV* pv = is->pv;
  int nvc = is->ncontacts;
  for (int k = 0; k < nvc; ++k) {
    VC* vc = &is->vc[k];
    PC* pc = &is->pc[k];
    Contact* c = &w->c[vc->contact];
    double ma = vc->ma, mb = vc->mb, Ia = vc->Ia, Ib = vc->Ib;
    V2 ca = pv[vc->ia].c, cb = pv[vc->ib].c;
    double aa = pv[vc->ia].a, ab = pv[vc->ib].a;
    V2 va = pv[vc->ia].v, vb = pv[vc->ib].v;
    double wa = pv[vc->ia].w, wb = pv[vc->ib].w;
    XF xfa = xf_from(ca, aa, pc->lca);
    XF xfb = xf_from(cb, ab, pc->lcb);
    V2 wpts[2];
    vc->normal = world_manifold(&c->m, xfa, pc->ra, xfb, pc->rb, wpts);
    for (int j = 0; j < vc->npts; ++j) {
      VCP* p = &vc->p[j];
      p->ra = sub(wpts[j], ca);
      p->rb = sub(wpts[j], cb);
      double rna = pysq(cross(p->ra, vc->normal));
      double rnb = pysq(cross(p->rb, vc->normal));
      double kn = ma + mb + Ia * rna + Ib * rnb;
      p->nm = kn > 0.0 ? 1.0 / kn : 0.0;
      V2 tangent = cross_vs(vc->normal, 1.0);
      double rta = pysq(cross(p->ra, tangent));
      double rtb = pysq(cross(p->rb, tangent));
      double kt = ma + mb + Ia * rta + Ib * rtb;
      p->tm = kt > 0.0 ? 1.0 / kt : 0.0;
      p->bias = 0.0;
      double vrel = dot(vc->normal, sub(sub(add(vb, cross_sv(wb, p->rb)), va), cross_sv(wa, p->ra)));
      if (vrel < -VELOCITY_THRESHOLD) p->bias = -vc->restitution * vrel;
    }
    if (vc->npts == 2) {
      V2 ra1 = vc->p[0].ra, rb1 = vc->p[0].rb, ra2 = vc->p[1].ra, rb2 = vc->p[1].rb;
      double rn1a = cross(ra1, vc->normal), rn1b = cross(rb1, vc->normal);
      double rn2a = cross(ra2, vc->normal), rn2b = cross(rb2, vc->normal);
      double k11 = ma + mb + Ia * rn1a * rn1a + Ib * rn1b * rn1b;
      double k22 = ma + mb + Ia * rn2a * rn2a + Ib * rn2b * rn2b;
      double k12 = ma + mb + Ia * rn1a * rn2a + Ib * rn1b * rn2b;
      if (pysq(k11) < (1000.0 * (k11 * k22 - pysq(k12)))) {
        vc->K.c1x = k11; vc->K.c1y = k12; vc->K.c2x = k12; vc->K.c2y = k22;
        /* PyMat22.inverse (common.py:502-515) */
        double A_ = vc->K.c1x, B_ = vc->K.c2x, C_ = vc->K.c1y, D_ = vc->K.c2y;
        double det = A_ * D_ - B_ * C_;
        if (det != 0.0) det = 1.0 / det;
        vc->nmass.c1x = det * D_; vc->nmass.c1y = -det * C_;
        vc->nmass.c2x = -det * B_; vc->nmass.c2y = det * A_;
      } else {
        vc->npts = 1; /* quirk Q6 */
      }
    }
  }
}

/* Island.report (island.py:245-252): post_solve with the velocity constraints' impulses, keyed
   from key0 in island order */
static void cs_report(World* w, IslandBuf* is, int phase, int key0) {
  for (int k = 0; k < is->ncontacts; ++k) {
    const VC* vc = &is->vc[k];
    int e = ev_push(w, B2G_EV_POST_SOLVE, phase, key0 + k, &w->c[vc->contact]);
    if (e < 0) continue;
    int32_t* ei = w->ev_i + (size_t)e * B2G_EV_NI;
    double* ed = w->ev_d + (size_t)e * B2G_EV_ND;
    ei[B2G_EI_OLD_PC] = vc->npts; /* the impulses handed to the listener */
    ed[B2G_ED_OLD + B2G_ED_P0_NI] = vc->p[0].ni;
    ed[B2G_ED_OLD + B2G_ED_P0_TI] = vc->p[0].ti;
    ed[B2G_ED_OLD + B2G_ED_P1_NI] = vc->npts >= 2 ? vc->p[1].ni : 0.0;
    ed[B2G_ED_OLD + B2G_ED_P1_TI] = vc->npts >= 2 ? vc->p[1].ti : 0.0;
  }
}

/* ------------------------------------------------------------------ */
/* Joints (pypybox2d/joints/): velocity-constraint setup, velocity and  */
/* position solves, called by Island.solve (island.py:107-148).  Each  */
/* joint's numbers live in its record (b2g_layout.h B2G_JD_*), which   */
/* stands for the joint object's attributes: values the reference      */
/* leaves stale between steps stay stale here too.                     */
/* ------------------------------------------------------------------ */
typedef struct { double x, y, z; } V3;
typedef struct { V3 c1, c2, c3; } M33;
static inline V3 v3(double x, double y, double z) { V3 r; r.x = x; r.y = y; r.z = z; return r; }
static inline V3 add3(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V3 sub3(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V3 neg3(V3 a) { return v3(-a.x, -a.y, -a.z); }
static inline V3 scale3(double s, V3 a) { return v3(s * a.x, s * a.y, s * a.z); }
static inline double dot3(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; } /* common.py:342-344 */
static inline V3 cross3(V3 s, V3 v) {                                              /* common.py:353-358 */
  return v3(s.y * v.z - s.z * v.y, s.z * v.x - s.x * v.z, s.x * v.y - s.y * v.x);
}
static V3 solve33(M33 m, V3 b) { /* Mat33.solve3x3 (common.py:587-599) */
  double det = dot3(m.c1, cross3(m.c2, m.c3));
  if (det != 0.0) det = 1.0 / det;
  return v3(det * dot3(b, cross3(m.c2, m.c3)), det * dot3(m.c1, cross3(b, m.c3)), det * dot3(m.c1, cross3(m.c2, b)));
}
static V2 solve22(M33 m, V2 v) { /* Mat33.solve2x2 (common.py:601-614) */
  double a11 = m.c1.x, a12 = m.c2.x, a21 = m.c1.y, a22 = m.c2.y;
  double det = a11 * a22 - a12 * a21;
  if (det != 0.0) det = 1.0 / det;
  return v2(det * (a22 * v.x - a12 * v.y), det * (a11 * v.y - a21 * v.x));
}
static V2 m22_solve(M22 m, V2 v) { /* Mat22.solve (common.py:521-531) */
  double a = m.c1x, b = m.c2x, c = m.c1y, d = m.c2y;
  double det = a * d - b * c;
  if (det != 0.0) det = 1.0 / det;
  return v2(det * (d * v.x - b * v.y), det * (a * v.y - c * v.x));
}
static M22 m22_inverse(M22 m) { /* Mat22.inverse (common.py:506-519) */
  double a = m.c1x, b = m.c2x, c = m.c1y, d = m.c2y;
  double det = a * d - b * c;
  if (det != 0.0) det = 1.0 / det;
  M22 r;
  r.c1x = det * d; r.c1y = -det * c;
  r.c2x = -det * b; r.c2y = det * a;
  return r;
}
static inline V2 vdiv(V2 a, double s) { return v2(a.x / s, a.y / s); } /* Vec2.__div__ */

static inline V2 jget2(const Joint* J, int k) { return v2(J->d[k], J->d[k + 1]); }
static inline void jset2(Joint* J, int k, V2 v) { J->d[k] = v.x; J->d[k + 1] = v.y; }
static inline V3 jget3(const Joint* J, int k) { return v3(J->d[k], J->d[k + 1], J->d[k + 2]); }
static inline void jset3(Joint* J, int k, V3 v) { J->d[k] = v.x; J->d[k + 1] = v.y; J->d[k + 2] = v.z; }
static inline M33 jget33(const Joint* J, int k) {
  M33 m; m.c1 = jget3(J, k); m.c2 = jget3(J, k + 3); m.c3 = jget3(J, k + 6); return m;
}
static inline void jset33(Joint* J, int k, M33 m) { jset3(J, k, m.c1); jset3(J, k + 3, m.c2); jset3(J, k + 6, m.c3); }
static inline M22 jget22(const Joint* J, int k) {
  M22 m; m.c1x = J->d[k]; m.c1y = J->d[k + 1]; m.c2x = J->d[k + 2]; m.c2y = J->d[k + 3]; return m;
}
static inline void jset22(Joint* J, int k, M22 m) {
  J->d[k] = m.c1x; J->d[k + 1] = m.c1y; J->d[k + 2] = m.c2x; J->d[k + 3] = m.c2y;
}

/* the two-body joints' common view of their bodies */
typedef struct {
  int ia, ib;               /* island indices */
  double ma, mb, Ia, Ib;    /* inverse masses and inertias */
  V2 lca, lcb;              /* local centres */
} JB;
static JB joint_bodies(const World* w, const Joint* J) {
  JB b;
  const Body* A = &w->b[J->i[B2G_JI_BODY_A]];
  const Body* B = &w->b[J->i[B2G_JI_BODY_B]];
  b.ia = A->island_index; b.ib = B->island_index;
  b.ma = A->inv_mass; b.mb = B->inv_mass;
  b.Ia = A->invI; b.Ib = B->invI;
  b.lca = A->lc; b.lcb = B->lc;
  return b;
}
#define JFLAG(J, f) (((J)->i[B2G_JI_FLAGS] & (f)) != 0)

/* ---- revolute (joints/revolute.py:214-453) ---- */
static void revolute_init(World* w, Joint* J, PV* pv, const Step* st) {
  JB b = joint_bodies(w, J);
  double ma = b.ma, mb = b.mb, ia = b.Ia, ib = b.Ib;
  double aa = pv[b.ia].a, ab = pv[b.ib].a;
  V2 va = pv[b.ia].v, vb = pv[b.ib].v;
  double wa = pv[b.ia].w, wb = pv[b.ib].w;
  M22 qa = rot(aa), qb = rot(ab);
  V2 ra = mulmv(qa, sub(jget2(J, B2G_JD_LA_X), b.lca));
  V2 rb = mulmv(qb, sub(jget2(J, B2G_JD_LB_X), b.lcb));
  jset2(J, B2G_JD_RA_X, ra); jset2(J, B2G_JD_RB_X, rb);
  M33 m;
  m.c3 = v3(-ra.y * ia - rb.y * ib, ra.x * ia + rb.x * ib, ia + ib);
  m.c2 = v3(-ra.y * ra.x * ia - rb.y * rb.x * ib, ma + mb + ra.x * ra.x * ia + rb.x * rb.x * ib, m.c3.y);
  m.c1 = v3(ma + mb + ra.y * ra.y * ia + rb.y * rb.y * ib, m.c2.x, m.c3.x);
  jset33(J, B2G_JD_REV_MASS, m);
  double motor_mass = ia + ib;
  int fixed_rotation = (motor_mass == 0.0);
  if (motor_mass > 0.0) motor_mass = 1.0 / motor_mass;
  J->d[B2G_JD_REV_MOTOR_MASS] = motor_mass;
  if (!JFLAG(J, B2G_JF_MOTOR) || fixed_rotation) J->d[B2G_JD_MOTOR_IMP] = 0.0;
  double* imp_z = &J->d[B2G_JD_IMP_Z];
  int* state = &J->i[B2G_JI_STATE];
  if (JFLAG(J, B2G_JF_LIMIT) && !fixed_rotation) {
    double joint_angle = ab - aa - J->d[B2G_JD_REF];
    double lower = J->d[B2G_JD_LOWER], upper = J->d[B2G_JD_UPPER];
    if (fabs(upper - lower) < 2.0 * ANGULAR_SLOP) {
      *state = B2G_LIMIT_EQUAL;
    } else if (joint_angle <= lower) {
      if (*state != B2G_LIMIT_AT_LOWER) *imp_z = *imp_z - *imp_z;
      *state = B2G_LIMIT_AT_LOWER;
    } else if (joint_angle >= upper) {
      if (*state != B2G_LIMIT_AT_UPPER) *imp_z = *imp_z - *imp_z;
      *state = B2G_LIMIT_AT_UPPER;
    } else {
      *state = B2G_LIMIT_INACTIVE;
      *imp_z = *imp_z - *imp_z;
    }
  } else {
    *state = B2G_LIMIT_INACTIVE;
  }
  if (st->warm_starting) {
    V3 imp = scale3(st->dt_ratio, jget3(J, B2G_JD_IMP_X));
    jset3(J, B2G_JD_IMP_X, imp);
    double mi = J->d[B2G_JD_MOTOR_IMP] * st->dt_ratio;
    J->d[B2G_JD_MOTOR_IMP] = mi;
    V2 P = v2(imp.x, imp.y);
    va = sub(va, scale(ma, P));
    wa -= ia * (cross(ra, P) + mi + imp.z);
    vb = add(vb, scale(mb, P));
    wb += ib * (cross(rb, P) + mi + imp.z);
  } else {
    jset3(J, B2G_JD_IMP_X, v3(0.0, 0.0, 0.0));
    J->d[B2G_JD_MOTOR_IMP] = 0.0;
  }
  pv[b.ia].v = va; pv[b.ia].w = wa;
  pv[b.ib].v = vb; pv[b.ib].w = wb;
}

static void revolute_solve_velocity(World* w, Joint* J, PV* pv, const Step* st) {
  JB b = joint_bodies(w, J);
  V2 va = pv[b.ia].v, vb = pv[b.ib].v;
  double wa = pv[b.ia].w, wb = pv[b.ib].w;
  double ma = b.ma, mb = b.mb, ia = b.Ia, ib = b.Ib;
  V2 ra = jget2(J, B2G_JD_RA_X), rb = jget2(J, B2G_JD_RB_X);
  M33 mass = jget33(J, B2G_JD_REV_MASS);
  int fixed_rotation = ((ia + ib) == 0.0);
  int state = J->i[B2G_JI_STATE];
  if (JFLAG(J, B2G_JF_MOTOR) && state != B2G_LIMIT_EQUAL && !fixed_rotation) {
    double Cdot = wb - wa - J->d[B2G_JD_MOTOR_SPEED];
    double impulse = J->d[B2G_JD_REV_MOTOR_MASS] * (-Cdot);
    double old = J->d[B2G_JD_MOTOR_IMP];
    double max_impulse = st->dt * J->d[B2G_JD_MAX_MOTOR];
    J->d[B2G_JD_MOTOR_IMP] = pclamp(old + impulse, -max_impulse, max_impulse);
    impulse = J->d[B2G_JD_MOTOR_IMP] - old;
    wa -= ia * impulse;
    wb += ib * impulse;
  }
  V3 acc = jget3(J, B2G_JD_IMP_X);
  if (JFLAG(J, B2G_JF_LIMIT) && state != B2G_LIMIT_INACTIVE && !fixed_rotation) {
    V2 Cdot1 = sub(sub(add(vb, cross_sv(wb, rb)), va), cross_sv(wa, ra));
    double Cdot2 = wb - wa;
    V3 Cdot = v3(Cdot1.x, Cdot1.y, Cdot2);
    V3 impulse = neg3(solve33(mass, Cdot));
    if (state == B2G_LIMIT_EQUAL) {
      acc = add3(acc, impulse);
    } else if (state == B2G_LIMIT_AT_LOWER) {
      double new_impulse = acc.z + impulse.z;
      if (new_impulse < 0.0) {
        V2 rhs = add(neg(Cdot1), scale(acc.z, v2(mass.c3.x, mass.c3.y)));
        V2 reduced = solve22(mass, rhs);
        impulse = v3(reduced.x, reduced.y, -acc.z);
        acc = add3(acc, v3(reduced.x, reduced.y, -acc.z));
      } else {
        acc = add3(acc, impulse);
      }
    } else if (state == B2G_LIMIT_AT_UPPER) {
      double new_impulse = acc.z + impulse.z;
      if (new_impulse > 0.0) {
        V2 rhs = add(neg(Cdot1), scale(acc.z, v2(mass.c3.x, mass.c3.y)));
        V2 reduced = solve22(mass, rhs);
        impulse = v3(reduced.x, reduced.y, -acc.z);
        acc = add3(acc, v3(reduced.x, reduced.y, -acc.z));
      } else {
        acc = add3(acc, impulse);
      }
    }
    V2 P = v2(impulse.x, impulse.y);
    va = sub(va, scale(ma, P));
    wa -= ia * (cross(ra, P) + impulse.z);
    vb = add(vb, scale(mb, P));
    wb += ib * (cross(rb, P) + impulse.z);
  } else {
    V2 Cdot = sub(sub(add(vb, cross_sv(wb, rb)), va), cross_sv(wa, ra));
    V2 impulse = solve22(mass, neg(Cdot));
    acc = add3(acc, v3(impulse.x, impulse.y, 0.0));
    va = sub(va, scale(ma, impulse));
    wa -= ia * cross(ra, impulse);
    vb = add(vb, scale(mb, impulse));
    wb += ib * cross(rb, impulse);
  }
  jset3(J, B2G_JD_IMP_X, acc);
  pv[b.ia].v = va; pv[b.ia].w = wa;
  pv[b.ib].v = vb; pv[b.ib].w = wb;
}

static int revolute_solve_position(World* w, Joint* J, PV* pv) {
  JB b = joint_bodies(w, J);
  V2 ca = pv[b.ia].c, cb = pv[b.ib].c;
  double aa = pv[b.ia].a, ab = pv[b.ib].a;
  double ma = b.ma, mb = b.mb, ia = b.Ia, ib = b.Ib;
  int fixed_rotation = ((ia + ib) == 0.0);
  double angular_error = 0.0, position_error = 0.0;
  int state = J->i[B2G_JI_STATE];
  if (JFLAG(J, B2G_JF_LIMIT) && state != B2G_LIMIT_INACTIVE && !fixed_rotation) {
    double angle = ab - aa - J->d[B2G_JD_REF];
    double limit_impulse = 0.0, C;
    double motor_mass = J->d[B2G_JD_REV_MOTOR_MASS];
    if (state == B2G_LIMIT_EQUAL) {
      C = pclamp(angle - J->d[B2G_JD_LOWER], -MAX_ANGULAR_CORRECTION, MAX_ANGULAR_CORRECTION);
      limit_impulse = -motor_mass * C;
      angular_error = fabs(C);
    } else if (state == B2G_LIMIT_AT_LOWER) {
      C = angle - J->d[B2G_JD_LOWER];
      angular_error = -C;
      C = pclamp(C + ANGULAR_SLOP, -MAX_ANGULAR_CORRECTION, 0.0);
      limit_impulse = -motor_mass * C;
    } else if (state == B2G_LIMIT_AT_UPPER) {
      C = angle - J->d[B2G_JD_UPPER];
      angular_error = C;
      C = pclamp(C - ANGULAR_SLOP, 0.0, MAX_ANGULAR_CORRECTION);
      limit_impulse = -motor_mass * C;
    }
    aa -= ia * limit_impulse;
    ab += ib * limit_impulse;
  }
  M22 qa = rot(aa), qb = rot(ab);
  V2 ra = mulmv(qa, sub(jget2(J, B2G_JD_LA_X), b.lca));
  V2 rb = mulmv(qb, sub(jget2(J, B2G_JD_LB_X), b.lcb));
  V2 C = sub(sub(add(cb, rb), ca), ra);
  position_error = length_(C);
  double rax = ra.x, ray = ra.y, rbx = rb.x, rby = rb.y;
  M22 K;
  K.c1x = ma + mb + ia * ray * ray + ib * rby * rby;
  K.c1y = -ia * rax * ray - ib * rbx * rby;
  K.c2x = K.c1y;
  K.c2y = ma + mb + ia * rax * rax + ib * rbx * rbx;
  V2 impulse = neg(m22_solve(K, C));
  ca = sub(ca, scale(ma, impulse));
  aa -= ia * cross(ra, impulse);
  cb = add(cb, scale(mb, impulse));
  ab += ib * cross(rb, impulse);
  pv[b.ia].c = ca; pv[b.ia].a = aa;
  pv[b.ib].c = cb; pv[b.ib].a = ab;
  return position_error <= LINEAR_SLOP && angular_error <= ANGULAR_SLOP;
}

/* ---- prismatic (joints/prismatic.py:271-513) ---- */
static M33 prismatic_k(double ma, double mb, double ia, double ib, double s1, double s2, double a1, double a2) {
  double k11 = ma + mb + ia * pysq(s1) + ib * pysq(s2);
  double k12 = ia * s1 + ib * s2;
  double k13 = ia * s1 * a1 + ib * s2 * a2;
  double k22 = ia + ib;
  if (k22 == 0.0) k22 = 1.0;
  double k23 = ia * a1 + ib * a2;
  double k33 = ma + mb + ia * pysq(a1) + ib * pysq(a2);
  M33 k;
  k.c1 = v3(k11, k12, k13);
  k.c2 = v3(k12, k22, k23);
  k.c3 = v3(k13, k23, k33);
  return k;
}

static void prismatic_init(World* w, Joint* J, PV* pv, const Step* st) {
  JB b = joint_bodies(w, J);
  double ma = b.ma, mb = b.mb, ia = b.Ia, ib = b.Ib;
  V2 ca = pv[b.ia].c, cb = pv[b.ib].c;
  double aa = pv[b.ia].a, ab = pv[b.ib].a;
  V2 va = pv[b.ia].v, vb = pv[b.ib].v;
  double wa = pv[b.ia].w, wb = pv[b.ib].w;
  M22 qa = rot(aa), qb = rot(ab);
  V2 ra = mulmv(qa, sub(jget2(J, B2G_JD_LA_X), b.lca));
  V2 rb = mulmv(qb, sub(jget2(J, B2G_JD_LB_X), b.lcb));
  jset2(J, B2G_JD_RA_X, ra); jset2(J, B2G_JD_RB_X, rb);
  V2 d = sub(sub(add(cb, rb), ca), ra);
  V2 axis = mulmv(qa, jget2(J, B2G_JD_AXIS_X));
  jset2(J, B2G_JD_PRI_AXIS_X, axis);
  double a1 = cross(add(d, ra), axis), a2 = cross(rb, axis);
  J->d[B2G_JD_PRI_A1] = a1; J->d[B2G_JD_PRI_A2] = a2;
  double motor_mass = ma + mb + ia * pysq(a1) + ib * pysq(a2);
  if (motor_mass > 0.0) motor_mass = 1.0 / motor_mass;
  J->d[B2G_JD_PRI_MOTOR_MASS] = motor_mass;
  V2 perp = mulmv(qa, jget2(J, B2G_JD_PERP_X));
  jset2(J, B2G_JD_PRI_PERP_X, perp);
  double s1 = cross(add(d, ra), perp), s2 = cross(rb, perp);
  J->d[B2G_JD_PRI_S1] = s1; J->d[B2G_JD_PRI_S2] = s2;
  jset33(J, B2G_JD_PRI_K, prismatic_k(ma, mb, ia, ib, s1, s2, a1, a2));
  double* imp_z = &J->d[B2G_JD_IMP_Z];
  int* state = &J->i[B2G_JI_STATE];
  if (JFLAG(J, B2G_JF_LIMIT)) {
    double jt = dot(axis, d);
    double lower = J->d[B2G_JD_LOWER], upper = J->d[B2G_JD_UPPER];
    if (fabs(upper - lower) < 2.0 * LINEAR_SLOP) {
      *state = B2G_LIMIT_EQUAL;
    } else if (jt <= lower) {
      if (*state != B2G_LIMIT_AT_LOWER) { *state = B2G_LIMIT_AT_LOWER; *imp_z = *imp_z - *imp_z; }
    } else if (jt >= upper) {
      if (*state != B2G_LIMIT_AT_UPPER) { *state = B2G_LIMIT_AT_UPPER; *imp_z = *imp_z - *imp_z; }
    } else {
      *state = B2G_LIMIT_INACTIVE;
      *imp_z = *imp_z - *imp_z;
    }
  } else {
    *state = B2G_LIMIT_INACTIVE;
    *imp_z = *imp_z - *imp_z;
  }
  if (!JFLAG(J, B2G_JF_MOTOR)) J->d[B2G_JD_MOTOR_IMP] = 0.0;
  if (st->warm_starting) {
    V3 imp = scale3(st->dt_ratio, jget3(J, B2G_JD_IMP_X));
    jset3(J, B2G_JD_IMP_X, imp);
    double mi = J->d[B2G_JD_MOTOR_IMP] * st->dt_ratio;
    J->d[B2G_JD_MOTOR_IMP] = mi;
    V2 P = add(scale(imp.x, perp), scale(mi + imp.z, axis));
    double L1 = imp.x * s1 + imp.y + (mi + imp.z) * a1;
    double L2 = imp.x * s2 + imp.y + (mi + imp.z) * a2;
    va = sub(va, scale(ma, P));
    wa -= ia * L1;
    vb = add(vb, scale(mb, P));
    wb += ib * L2;
  } else {
    jset3(J, B2G_JD_IMP_X, v3(0.0, 0.0, 0.0));
    J->d[B2G_JD_MOTOR_IMP] = 0.0;
  }
  pv[b.ia].v = va; pv[b.ia].w = wa;
  pv[b.ib].v = vb; pv[b.ib].w = wb;
}

static void prismatic_solve_velocity(World* w, Joint* J, PV* pv, const Step* st) {
  JB b = joint_bodies(w, J);
  V2 va = pv[b.ia].v, vb = pv[b.ib].v;
  double wa = pv[b.ia].w, wb = pv[b.ib].w;
  double ma = b.ma, mb = b.mb, ia = b.Ia, ib = b.Ib;
  double a1 = J->d[B2G_JD_PRI_A1], a2 = J->d[B2G_JD_PRI_A2];
  double s1 = J->d[B2G_JD_PRI_S1], s2 = J->d[B2G_JD_PRI_S2];
  V2 perp = jget2(J, B2G_JD_PRI_PERP_X), axis = jget2(J, B2G_JD_PRI_AXIS_X);
  M33 k = jget33(J, B2G_JD_PRI_K);
  int state = J->i[B2G_JI_STATE];
  if (JFLAG(J, B2G_JF_MOTOR) && state != B2G_LIMIT_EQUAL) {
    double Cdot = dot(axis, sub(vb, va)) + a2 * wb - a1 * wa;
    double impulse = J->d[B2G_JD_PRI_MOTOR_MASS] * (J->d[B2G_JD_MOTOR_SPEED] - Cdot);
    double old = J->d[B2G_JD_MOTOR_IMP];
    double max_impulse = st->dt * J->d[B2G_JD_MAX_MOTOR];
    J->d[B2G_JD_MOTOR_IMP] = pclamp(old + impulse, -max_impulse, max_impulse);
    impulse = J->d[B2G_JD_MOTOR_IMP] - old;
    V2 P = scale(impulse, axis);
    double L1 = impulse * a1, L2 = impulse * a2;
    va = sub(va, scale(ma, P));
    wa -= ia * L1;
    vb = add(vb, scale(mb, P));
    wb += ib * L2;
  }
  V2 Cdot1 = v2(dot(perp, sub(vb, va)) + s2 * wb - s1 * wa, wb - wa);
  V2 P;
  double L1, L2;
  V3 acc = jget3(J, B2G_JD_IMP_X);
  if (JFLAG(J, B2G_JF_LIMIT) && state != B2G_LIMIT_INACTIVE) {
    double Cdot2 = dot(axis, sub(vb, va)) + a2 * wb - a1 * wa;
    V3 Cdot = v3(Cdot1.x, Cdot1.y, Cdot2);
    V3 old = acc;
    V3 df = solve33(k, neg3(Cdot));
    acc = add3(acc, df);
    double iz;
    if (state == B2G_LIMIT_AT_LOWER) iz = pmax(acc.z, 0.0);
    else if (state == B2G_LIMIT_AT_UPPER) iz = pmin(acc.z, 0.0);
    else iz = acc.z;
    V2 bb = sub(neg(Cdot1), scale(acc.z - old.z, v2(k.c3.x, k.c3.y)));
    V2 f2r = add(solve22(k, bb), v2(old.x, old.y));
    acc = v3(f2r.x, f2r.y, iz);
    df = sub3(acc, old);
    P = add(scale(df.x, perp), scale(df.z, axis));
    L1 = df.x * s1 + df.y + df.z * a1;
    L2 = df.x * s2 + df.y + df.z * a2;
  } else {
    V2 df = solve22(k, neg(Cdot1));
    acc = add3(acc, v3(df.x, df.y, 0.0));
    P = scale(df.x, perp);
    L1 = df.x * s1 + df.y;
    L2 = df.x * s2 + df.y;
  }
  jset3(J, B2G_JD_IMP_X, acc);
  va = sub(va, scale(ma, P));
  wa -= ia * L1;
  vb = add(vb, scale(mb, P));
  wb += ib * L2;
  pv[b.ia].v = va; pv[b.ia].w = wa;
  pv[b.ib].v = vb; pv[b.ib].w = wb;
}

static int prismatic_solve_position(World* w, Joint* J, PV* pv) {
  JB b = joint_bodies(w, J);
  V2 ca = pv[b.ia].c, cb = pv[b.ib].c;
  double aa = pv[b.ia].a, ab = pv[b.ib].a;
  double ma = b.ma, mb = b.mb, ia = b.Ia, ib = b.Ib;
  M22 qa = rot(aa), qb = rot(ab);
  V2 ra = mulmv(qa, sub(jget2(J, B2G_JD_LA_X), b.lca));
  V2 rb = mulmv(qb, sub(jget2(J, B2G_JD_LB_X), b.lcb));
  V2 d = sub(sub(add(cb, rb), ca), ra);
  V2 axis = mulmv(qa, jget2(J, B2G_JD_AXIS_X));
  double a1 = cross(add(d, ra), axis), a2 = cross(rb, axis);
  V2 perp = mulmv(qa, jget2(J, B2G_JD_PERP_X));
  double s1 = cross(add(d, ra), perp), s2 = cross(rb, perp);
  V2 C1 = v2(dot(perp, d), ab - aa - J->d[B2G_JD_REF]);
  double linear_error = fabs(C1.x), angular_error = fabs(C1.y);
  int active = 0;
  double C2 = 0.0;
  if (JFLAG(J, B2G_JF_LIMIT)) {
    double translation = dot(axis, d);
    double lower = J->d[B2G_JD_LOWER], upper = J->d[B2G_JD_UPPER];
    if (fabs(upper - lower) < 2.0 * LINEAR_SLOP) {
      C2 = pclamp(translation, -MAX_LINEAR_CORRECTION, MAX_LINEAR_CORRECTION);
      linear_error = pmax(linear_error, translation);
      active = 1;
    } else if (translation <= lower) {
      C2 = pclamp(translation - lower + LINEAR_SLOP, -MAX_LINEAR_CORRECTION, 0.0);
      linear_error = pmax(linear_error, lower - translation);
      active = 1;
    } else if (translation >= upper) {
      C2 = pclamp(translation - upper - LINEAR_SLOP, 0.0, MAX_LINEAR_CORRECTION);
      linear_error = pmax(linear_error, translation - upper);
      active = 1;
    }
  }
  V3 impulse;
  if (active) {
    M33 k = prismatic_k(ma, mb, ia, ib, s1, s2, a1, a2);
    jset33(J, B2G_JD_PRI_K, k);
    impulse = solve33(k, neg3(v3(C1.x, C1.y, C2)));
  } else {
    double k11 = ma + mb + ia * s1 * s1 + ib * s2 * s2;
    double k12 = ia * s1 + ib * s2;
    double k22 = ia + ib;
    if (k22 == 0.0) k22 = 1.0;
    M33 k = jget33(J, B2G_JD_PRI_K);
    k.c1 = v3(k11, k12, 0.0);
    k.c2 = v3(k12, k22, 0.0);
    jset33(J, B2G_JD_PRI_K, k);
    V2 impulse1 = solve22(k, neg(C1));
    impulse = v3(impulse1.x, impulse1.y, 0.0);
  }
  V2 P = add(scale(impulse.x, perp), scale(impulse.z, axis));
  double L1 = impulse.x * s1 + impulse.y + impulse.z * a1;
  double L2 = impulse.x * s2 + impulse.y + impulse.z * a2;
  ca = sub(ca, scale(ma, P));
  cb = add(cb, scale(mb, P));
  aa -= ia * L1;
  ab += ib * L2;
  pv[b.ia].c = ca; pv[b.ia].a = aa;
  pv[b.ib].c = cb; pv[b.ib].a = ab;
  return (linear_error <= LINEAR_SLOP) && (angular_error <= ANGULAR_SLOP);
}

/* ---- distance (joints/distance.py:170-311); the effective mass uses the separation before
   normalisation: `self._u *= 1/length` rebinds the attribute, not the local u ---- */
static void distance_init(World* w, Joint* J, PV* pv, const Step* st) {
  JB b = joint_bodies(w, J);
  double ma = b.ma, mb = b.mb, ia = b.Ia, ib = b.Ib;
  V2 ca = pv[b.ia].c, cb = pv[b.ib].c;
  double aa = pv[b.ia].a, ab = pv[b.ib].a;
  V2 va = pv[b.ia].v, vb = pv[b.ib].v;
  double wa = pv[b.ia].w, wb = pv[b.ib].w;
  M22 qa = rot(aa), qb = rot(ab);
  V2 ra = mulmv(qa, sub(jget2(J, B2G_JD_LA_X), b.lca));
  V2 rb = mulmv(qb, sub(jget2(J, B2G_JD_LB_X), b.lcb));
  jset2(J, B2G_JD_RA_X, ra); jset2(J, B2G_JD_RB_X, rb);
  V2 u = sub(sub(add(cb, rb), ca), ra);
  double length = length_(u);
  V2 un = length > LINEAR_SLOP ? scale(1.0 / length, u) : v2(0.0, 0.0);
  jset2(J, B2G_JD_DIS_U_X, un);
  double cr_au = pysq(cross(ra, u)), cr_bu = pysq(cross(rb, u));
  double inv_mass = ma + ia * cr_au + mb + ib * cr_bu;
  double mass = inv_mass != 0.0 ? 1.0 / inv_mass : 0.0;
  J->d[B2G_JD_DIS_MASS] = mass;
  if (J->d[B2G_JD_FREQ] > 0.0) {
    double C = length - J->d[B2G_JD_LENGTH];
    double omega = 2.0 * PI_ * J->d[B2G_JD_FREQ];
    double dd = 2.0 * mass * J->d[B2G_JD_DAMP] * omega;
    double k = mass * pysq(omega);
    double dt = st->dt;
    double gamma = dt * (dd + dt * k);
    gamma = gamma != 0.0 ? 1.0 / gamma : 0.0;
    J->d[B2G_JD_DIS_GAMMA] = gamma;
    J->d[B2G_JD_DIS_BIAS] = C * dt * k * gamma;
    double m2 = inv_mass + gamma;
    J->d[B2G_JD_DIS_MASS] = m2 != 0.0 ? 1.0 / m2 : 0.0;
  }
  if (st->warm_starting) {
    double imp = J->d[B2G_JD_IMP_X] * st->dt_ratio;
    J->d[B2G_JD_IMP_X] = imp;
    V2 P = scale(imp, un);
    va = sub(va, scale(ma, P));
    wa -= ia * cross(ra, P);
    vb = add(vb, scale(mb, P));
    wb += ib * cross(rb, P);
  } else {
    J->d[B2G_JD_IMP_X] = 0.0;
  }
  pv[b.ia].v = va; pv[b.ia].w = wa;
  pv[b.ib].v = vb; pv[b.ib].w = wb;
}

static void distance_solve_velocity(World* w, Joint* J, PV* pv) {
  JB b = joint_bodies(w, J);
  V2 va = pv[b.ia].v, vb = pv[b.ib].v;
  double wa = pv[b.ia].w, wb = pv[b.ib].w;
  V2 ra = jget2(J, B2G_JD_RA_X), rb = jget2(J, B2G_JD_RB_X);
  V2 u = jget2(J, B2G_JD_DIS_U_X);
  V2 vpa = add(va, cross_sv(wa, ra)), vpb = add(vb, cross_sv(wb, rb));
  double Cdot = dot(u, sub(vpb, vpa));
  double impulse = -J->d[B2G_JD_DIS_MASS] * (Cdot + J->d[B2G_JD_DIS_BIAS] + J->d[B2G_JD_DIS_GAMMA] * J->d[B2G_JD_IMP_X]);
  J->d[B2G_JD_IMP_X] += impulse;
  V2 P = scale(impulse, u);
  va = sub(va, scale(b.ma, P));
  wa -= b.Ia * cross(ra, P);
  vb = add(vb, scale(b.mb, P));
  wb += b.Ib * cross(rb, P);
  pv[b.ia].v = va; pv[b.ia].w = wa;
  pv[b.ib].v = vb; pv[b.ib].w = wb;
}

static int distance_solve_position(World* w, Joint* J, PV* pv) {
  if (J->d[B2G_JD_FREQ] > 0.0) return 1; /* no position correction for soft constraints */
  JB b = joint_bodies(w, J);
  V2 ca = pv[b.ia].c, cb = pv[b.ib].c;
  double aa = pv[b.ia].a, ab = pv[b.ib].a;
  M22 qa = rot(aa), qb = rot(ab);
  V2 ra = mulmv(qa, sub(jget2(J, B2G_JD_LA_X), b.lca));
  V2 rb = mulmv(qb, sub(jget2(J, B2G_JD_LB_X), b.lcb));
  V2 u = sub(sub(add(cb, rb), ca), ra);
  double length = length_(u);
  u = normalized(u);
  double C = length - J->d[B2G_JD_LENGTH];
  double impulse = -J->d[B2G_JD_DIS_MASS] * pclamp(C, -MAX_LINEAR_CORRECTION, MAX_LINEAR_CORRECTION);
  V2 P = scale(impulse, u);
  ca = sub(ca, scale(b.ma, P));
  aa -= b.Ia * cross(ra, P);
  cb = add(cb, scale(b.mb, P));
  ab += b.Ib * cross(rb, P);
  pv[b.ia].c = ca; pv[b.ia].a = aa;
  pv[b.ib].c = cb; pv[b.ib].a = ab;
  return fabs(C) < LINEAR_SLOP;
}

/* ---- pulley (joints/pulley.py:181-320); the warm start and the position correction move the
   bodies by the effective masses (the local ma, mb are rebound before they are used) ---- */
static void pulley_init(World* w, Joint* J, PV* pv, const Step* st) {
  JB b = joint_bodies(w, J);
  double ma = b.ma, mb = b.mb, ia = b.Ia, ib = b.Ib;
  V2 ca = pv[b.ia].c, cb = pv[b.ib].c;
  double aa = pv[b.ia].a, ab = pv[b.ib].a;
  V2 va = pv[b.ia].v, vb = pv[b.ib].v;
  double wa = pv[b.ia].w, wb = pv[b.ib].w;
  M22 qa = rot(aa), qb = rot(ab);
  V2 ra = mulmv(qa, sub(jget2(J, B2G_JD_LA_X), b.lca));
  V2 rb = mulmv(qb, sub(jget2(J, B2G_JD_LB_X), b.lcb));
  jset2(J, B2G_JD_RA_X, ra); jset2(J, B2G_JD_RB_X, rb);
  V2 ua = sub(add(ca, ra), jget2(J, B2G_JD_G0_X));
  V2 ub = sub(add(cb, rb), jget2(J, B2G_JD_G1_X));
  double length_a = length_(ua), length_b = length_(ub);
  ua = length_a > 10.0 * LINEAR_SLOP ? vdiv(ua, length_a) : v2(0.0, 0.0);
  ub = length_b > 10.0 * LINEAR_SLOP ? vdiv(ub, length_b) : v2(0.0, 0.0);
  jset2(J, B2G_JD_PUL_UA_X, ua); jset2(J, B2G_JD_PUL_UB_X, ub);
  double rua = cross(ra, ua), rub = cross(rb, ub);
  ma = ma + ia * pysq(rua);
  mb = mb + ib * pysq(rub);
  double ratio = J->d[B2G_JD_RATIO];
  double mass = ma + ratio * ratio * mb;
  if (mass > 0.0) mass = 1.0 / mass;
  J->d[B2G_JD_PUL_MASS] = mass;
  if (st->warm_starting) {
    double imp = J->d[B2G_JD_IMP_X] * st->dt_ratio;
    J->d[B2G_JD_IMP_X] = imp;
    V2 Pa = scale(-imp, ua);
    V2 Pb = scale(-ratio * imp, ub);
    va = add(va, scale(ma, Pa));
    wa += ia * cross(ra, Pa);
    vb = add(vb, scale(mb, Pb));
    wb += ib * cross(rb, Pb);
  } else {
    J->d[B2G_JD_IMP_X] = 0.0;
  }
  pv[b.ia].v = va; pv[b.ia].w = wa;
  pv[b.ib].v = vb; pv[b.ib].w = wb;
}

static void pulley_solve_velocity(World* w, Joint* J, PV* pv) {
  JB b = joint_bodies(w, J);
  V2 va = pv[b.ia].v, vb = pv[b.ib].v;
  double wa = pv[b.ia].w, wb = pv[b.ib].w;
  V2 ra = jget2(J, B2G_JD_RA_X), rb = jget2(J, B2G_JD_RB_X);
  V2 ua = jget2(J, B2G_JD_PUL_UA_X), ub = jget2(J, B2G_JD_PUL_UB_X);
  double ratio = J->d[B2G_JD_RATIO];
  V2 vpa = add(va, cross_sv(wa, ra)), vpb = add(vb, cross_sv(wb, rb));
  double Cdot = -dot(ua, vpa) - ratio * dot(ub, vpb);
  double impulse = -J->d[B2G_JD_PUL_MASS] * Cdot;
  J->d[B2G_JD_IMP_X] += impulse;
  V2 Pa = scale(-impulse, ua);
  V2 Pb = scale(-ratio * impulse, ub);
  va = add(va, scale(b.ma, Pa));
  wa += b.Ia * cross(ra, Pa);
  vb = add(vb, scale(b.mb, Pb));
  wb += b.Ib * cross(rb, Pb);
  pv[b.ia].v = va; pv[b.ia].w = wa;
  pv[b.ib].v = vb; pv[b.ib].w = wb;
}

static int pulley_solve_position(World* w, Joint* J, PV* pv) {
  JB b = joint_bodies(w, J);
  V2 ca = pv[b.ia].c, cb = pv[b.ib].c;
  double aa = pv[b.ia].a, ab = pv[b.ib].a;
  double ma = b.ma, mb = b.mb, ia = b.Ia, ib = b.Ib;
  M22 qa = rot(aa), qb = rot(ab);
  V2 ra = mulmv(qa, sub(jget2(J, B2G_JD_LA_X), b.lca));
  V2 rb = mulmv(qb, sub(jget2(J, B2G_JD_LB_X), b.lcb));
  V2 ua = sub(add(ca, ra), jget2(J, B2G_JD_G0_X));
  V2 ub = sub(add(cb, rb), jget2(J, B2G_JD_G1_X));
  double length_a = length_(ua), length_b = length_(ub);
  ua = length_a > 10.0 * LINEAR_SLOP ? vdiv(ua, length_a) : v2(0.0, 0.0);
  ub = length_b > 10.0 * LINEAR_SLOP ? vdiv(ub, length_b) : v2(0.0, 0.0);
  double rua = cross(ra, ua), rub = cross(rb, ub);
  ma = ma + ia * pysq(rua);
  mb = mb + ib * pysq(rub);
  double ratio = J->d[B2G_JD_RATIO];
  double mass = ma + pysq(ratio) * mb;
  if (mass > 0.0) mass = 1.0 / mass;
  double C = J->d[B2G_JD_LENGTH] - length_a - ratio * length_b;
  double linear_error = fabs(C);
  double impulse = -mass * C;
  V2 Pa = scale(-impulse, ua);
  V2 Pb = scale(-ratio * impulse, ub);
  ca = add(ca, scale(ma, Pa));
  aa += ia * cross(ra, Pa);
  cb = add(cb, scale(mb, Pb));
  ab += ib * cross(rb, Pb);
  pv[b.ia].c = ca; pv[b.ia].a = aa;
  pv[b.ib].c = cb; pv[b.ib].a = ab;
  return linear_error < LINEAR_SLOP;
}

/* ---- mouse (joints/mouse.py:150-229): one body (body_b) ---- */
static void mouse_init(World* w, Joint* J, PV* pv, const Step* st) {
  const Body* B = &w->b[J->i[B2G_JI_BODY_B]];
  int ix = B->island_index;
  V2 cb = pv[ix].c;
  double ab = pv[ix].a;
  V2 vb = pv[ix].v;
  double wb = pv[ix].w;
  double mb = B->inv_mass, ib = B->invI;
  M22 qb = rot(ab);
  double mass = B->mass;
  double omega = 2.0 * PI_ * J->d[B2G_JD_FREQ];
  double d = 2.0 * mass * J->d[B2G_JD_DAMP] * omega;
  double k = mass * pysq(omega);
  double dt = st->dt;
  if (!(d + dt * k > EPSILON)) w->err |= ORC_E_JOINT; /* the reference asserts (mouse.py:173) */
  double gamma = dt * (d + dt * k);
  if (gamma != 0.0) gamma = 1.0 / gamma;
  double beta = dt * k * gamma;
  J->d[B2G_JD_MOU_GAMMA] = gamma;
  J->d[B2G_JD_MOU_BETA] = beta;
  V2 rb = mulmv(qb, sub(jget2(J, B2G_JD_LB_X), B->lc));
  jset2(J, B2G_JD_RB_X, rb);
  M22 K;
  K.c1x = mb + ib * pysq(rb.y) + gamma;
  K.c1y = -ib * rb.x * rb.y;
  K.c2x = K.c1y;
  K.c2y = mb + ib * pysq(rb.x) + gamma;
  jset22(J, B2G_JD_MOU_MASS, m22_inverse(K));
  jset2(J, B2G_JD_MOU_C_X, scale(beta, sub(add(cb, rb), jget2(J, B2G_JD_G0_X))));
  wb *= 0.98;
  if (st->warm_starting) {
    V2 imp = scale(st->dt_ratio, jget2(J, B2G_JD_IMP_X));
    jset2(J, B2G_JD_IMP_X, imp);
    vb = add(vb, scale(mb, imp));
    wb += ib * cross(rb, imp);
  } else {
    jset2(J, B2G_JD_IMP_X, v2(0.0, 0.0));
  }
  pv[ix].v = vb; pv[ix].w = wb;
}

static void mouse_solve_velocity(World* w, Joint* J, PV* pv, const Step* st) {
  const Body* B = &w->b[J->i[B2G_JI_BODY_B]];
  int ix = B->island_index;
  V2 vb = pv[ix].v;
  double wb = pv[ix].w;
  V2 rb = jget2(J, B2G_JD_RB_X);
  V2 Cdot = add(vb, cross_sv(wb, rb));
  V2 acc = jget2(J, B2G_JD_IMP_X);
  V2 impulse = mulmv(jget22(J, B2G_JD_MOU_MASS),
                     neg(add(add(Cdot, jget2(J, B2G_JD_MOU_C_X)), scale(J->d[B2G_JD_MOU_GAMMA], acc))));
  V2 old = acc;
  acc = add(acc, impulse);
  double max_impulse = st->dt * J->d[B2G_JD_MAX_MOTOR];
  if (length_sq(acc) > pysq(max_impulse)) acc = scale(max_impulse / length_(acc), acc);
  jset2(J, B2G_JD_IMP_X, acc);
  impulse = sub(acc, old);
  vb = add(vb, scale(B->inv_mass, impulse));
  wb += B->invI * cross(rb, impulse);
  pv[ix].v = vb; pv[ix].w = wb;
}

/* ---- gear (joints/gear.py:262-388).  The constructor takes bodies b and d from joint_a as
   well (gear.py:91-92), so a == b and c == d here; the velocity solve keeps the linear
   velocities it captured at setup (only the angular ones are re-read, gear.py:337-340), and the
   position solve starts from the positions captured at setup (gear.py:356-358) and writes them
   back with its correction. ---- */
typedef struct { int idx[4]; double m[4], I[4]; V2 lc[4]; } GB;
static GB gear_bodies(const World* w, const Joint* J) {
  static const int slot[4] = {B2G_JI_BODY_A, B2G_JI_BODY_B, B2G_JI_BODY_C, B2G_JI_BODY_D};
  GB g;
  for (int k = 0; k < 4; ++k) {
    const Body* b = &w->b[J->i[slot[k]]];
    g.idx[k] = b->island_index;
    g.m[k] = b->inv_mass;
    g.I[k] = b->invI;
    g.lc[k] = b->lc;
  }
  return g;
}

static void gear_init(World* w, Joint* J, PV* pv, const Step* st) {
  GB g = gear_bodies(w, J);
  V2 Cs[4], Vs[4];
  double As[4], Ws[4];
  for (int k = 0; k < 4; ++k) {
    Cs[k] = pv[g.idx[k]].c; As[k] = pv[g.idx[k]].a;
    Vs[k] = pv[g.idx[k]].v; Ws[k] = pv[g.idx[k]].w;
  }
  M22 qa = rot(As[0]), qb = rot(As[1]), qc = rot(As[2]), qd = rot(As[3]);
  double ia = g.I[0], ib = g.I[1], ic = g.I[2], id_ = g.I[3];
  double ma = g.m[0], mb = g.m[1], mc = g.m[2], md = g.m[3];
  double ratio = J->d[B2G_JD_RATIO];
  double mass = 0.0;
  V2 j_vac, j_vbd;
  double j_wa, j_wb, j_wc, j_wd;
  if (JFLAG(J, B2G_JF_REVOLUTE_A)) {
    j_vac = v2(0.0, 0.0);
    j_wa = 1.0;
    j_wc = 1.0;
    mass += ia + ic;
  } else {
    V2 u = mulmv(qc, jget2(J, B2G_JD_AXIS_X));
    V2 rc = mulmv(qc, sub(jget2(J, B2G_JD_G0_X), g.lc[2]));
    V2 ra = mulmv(qa, sub(jget2(J, B2G_JD_LA_X), g.lc[0]));
    j_vac = u;
    j_wc = cross(rc, u);
    j_wa = cross(ra, u);
    mass += mc + ma + ic * j_wc * j_wc + ia * j_wa * j_wa;
  }
  if (JFLAG(J, B2G_JF_REVOLUTE_B)) {
    j_vbd = v2(0.0, 0.0);
    j_wb = 1.0;
    j_wd = 1.0;
    mass += pysq(ratio) * (ib + id_);
  } else {
    V2 u = mulmv(qd, jget2(J, B2G_JD_PERP_X));
    V2 rd = mulmv(qd, sub(jget2(J, B2G_JD_G1_X), g.lc[3]));
    V2 rb = mulmv(qb, sub(jget2(J, B2G_JD_LB_X), g.lc[1]));
    j_vbd = u;
    j_wd = cross(rd, u);
    j_wb = cross(rb, u);
    mass += pysq(ratio) * (md + mb) + id_ * pysq(j_wd) + ib * j_wb * j_wb;
  }
  mass = mass > 0.0 ? 1.0 / mass : 0.0;
  jset2(J, B2G_JD_GEA_JVAC_X, j_vac); jset2(J, B2G_JD_GEA_JVBD_X, j_vbd);
  J->d[B2G_JD_GEA_JWA] = j_wa; J->d[B2G_JD_GEA_JWB] = j_wb;
  J->d[B2G_JD_GEA_JWC] = j_wc; J->d[B2G_JD_GEA_JWD] = j_wd;
  J->d[B2G_JD_GEA_MASS] = mass;
  if (st->warm_starting) { /* no dt_ratio scaling here (gear.py:319-328) */
    double impulse = J->d[B2G_JD_IMP_X];
    Vs[0] = add(Vs[0], scale(ma * impulse, j_vac));
    Ws[0] += ia * impulse * j_wa;
    Vs[1] = add(Vs[1], scale(mb * impulse, j_vbd));
    Ws[1] += ib * impulse * j_wb;
    Vs[2] = sub(Vs[2], scale(mc * impulse, j_vac));
    Ws[2] -= ic * impulse * j_wc;
    Vs[3] = sub(Vs[3], scale(md * impulse, j_vbd));
    Ws[3] -= id_ * impulse * j_wd;
  } else {
    J->d[B2G_JD_IMP_X] = 0.0;
  }
  for (int k = 0; k < 4; ++k) {
    jset2(J, B2G_JD_GEA_V + 2 * k, Vs[k]);
    jset2(J, B2G_JD_GEA_C + 2 * k, Cs[k]);
    J->d[B2G_JD_GEA_A + k] = As[k];
    pv[g.idx[k]].v = Vs[k];
    pv[g.idx[k]].w = Ws[k];
  }
}

static void gear_solve_velocity(World* w, Joint* J, PV* pv) {
  GB g = gear_bodies(w, J);
  V2 Vs[4];
  double Ws[4];
  for (int k = 0; k < 4; ++k) {
    Vs[k] = jget2(J, B2G_JD_GEA_V + 2 * k);
    Ws[k] = pv[g.idx[k]].w;
  }
  V2 j_vac = jget2(J, B2G_JD_GEA_JVAC_X), j_vbd = jget2(J, B2G_JD_GEA_JVBD_X);
  double j_wa = J->d[B2G_JD_GEA_JWA], j_wb = J->d[B2G_JD_GEA_JWB];
  double j_wc = J->d[B2G_JD_GEA_JWC], j_wd = J->d[B2G_JD_GEA_JWD];
  double Cdot = dot(j_vac, sub(Vs[0], Vs[2])) + dot(j_vbd, sub(Vs[1], Vs[3]));
  Cdot += (j_wa * Ws[0] - j_wc * Ws[2]) + (j_wb * Ws[1] - j_wd * Ws[3]);
  double impulse = -J->d[B2G_JD_GEA_MASS] * Cdot;
  J->d[B2G_JD_IMP_X] += impulse;
  V2 lvj[4] = {j_vac, j_vbd, neg(j_vac), neg(j_vbd)};
  double avj[4] = {j_wa, j_wb, -j_wc, -j_wd};
  for (int k = 0; k < 4; ++k) {
    Vs[k] = add(Vs[k], scale(g.m[k] * impulse, lvj[k]));
    Ws[k] += (g.I[k] * impulse) * avj[k];
  }
  for (int k = 0; k < 4; ++k) {
    jset2(J, B2G_JD_GEA_V + 2 * k, Vs[k]);
    pv[g.idx[k]].v = Vs[k];
    pv[g.idx[k]].w = Ws[k];
  }
}

static int gear_solve_position(World* w, Joint* J, PV* pv) {
  GB g = gear_bodies(w, J);
  double ratio = J->d[B2G_JD_RATIO];
  V2 ca = jget2(J, B2G_JD_GEA_C), cb = jget2(J, B2G_JD_GEA_C + 2);
  V2 cc = jget2(J, B2G_JD_GEA_C + 4), cd = jget2(J, B2G_JD_GEA_C + 6);
  double aa = J->d[B2G_JD_GEA_A], ab = J->d[B2G_JD_GEA_A + 1];
  double ac = J->d[B2G_JD_GEA_A + 2], ad = J->d[B2G_JD_GEA_A + 3];
  M22 qa = rot(aa), qb = rot(ab), qc = rot(ac), qd = rot(ad);
  double ia = g.I[0], ib = g.I[1], ic = g.I[2], id_ = g.I[3];
  double ma = g.m[0], mb = g.m[1], mc = g.m[2], md = g.m[3];
  double mass = 0.0, coordinate_a, coordinate_b;
  V2 j_vac, j_vbd;
  double j_wa, j_wb, j_wc, j_wd;
  if (JFLAG(J, B2G_JF_REVOLUTE_A)) {
    j_vac = v2(0.0, 0.0);
    j_wa = 1.0;
    j_wc = 1.0;
    mass += ia + ic;
    coordinate_a = aa - ac - J->d[B2G_JD_REF];
  } else {
    V2 axis_c = jget2(J, B2G_JD_AXIS_X);
    V2 u = mulmv(qc, axis_c);
    V2 rc = mulmv(qc, sub(jget2(J, B2G_JD_G0_X), g.lc[2]));
    V2 ra = mulmv(qa, sub(jget2(J, B2G_JD_LA_X), g.lc[0]));
    j_vac = u;
    j_wc = cross(rc, u);
    j_wa = cross(ra, u);
    mass += mc + ma + ic * pysq(j_wc) + ia * pysq(j_wa);
    V2 pc = sub(jget2(J, B2G_JD_G0_X), g.lc[2]);
    V2 pa = mulmtv(qc, add(ra, sub(ca, cc)));
    coordinate_a = dot(sub(pa, pc), axis_c);
  }
  if (JFLAG(J, B2G_JF_REVOLUTE_B)) {
    j_vbd = v2(0.0, 0.0);
    j_wb = 1.0;
    j_wd = 1.0;
    mass += ib + id_;
    coordinate_b = ab - ad - J->d[B2G_JD_REF_B];
  } else {
    V2 axis_d = jget2(J, B2G_JD_PERP_X);
    V2 u = mulmv(qd, axis_d);
    V2 rd = mulmv(qd, sub(jget2(J, B2G_JD_G1_X), g.lc[3]));
    V2 rb = mulmv(qb, sub(jget2(J, B2G_JD_LB_X), g.lc[1]));
    j_vbd = scale(ratio, u);
    j_wd = ratio * cross(rd, u);
    j_wb = ratio * cross(rb, u);
    mass += pysq(ratio) * (md + mb) + id_ * pysq(j_wd) + ib * pysq(j_wb);
    V2 pd = sub(jget2(J, B2G_JD_G1_X), g.lc[3]);
    V2 pb = mulmtv(qd, add(rb, sub(cb, cd)));
    coordinate_b = dot(sub(pb, pd), axis_d);
  }
  double C = (coordinate_a + ratio * coordinate_b) - J->d[B2G_JD_LENGTH];
  if (mass > 0.0) {
    double impulse = -C / mass;
    ca = add(ca, scale(ma * impulse, j_vac));
    cb = add(cb, scale(mb * impulse, j_vbd));
    cc = sub(cc, scale(mc * impulse, j_vac));
    cd = sub(cd, scale(md * impulse, j_vbd));
    aa += ia * impulse * j_wa;
    ab += ib * impulse * j_wb;
    ac -= ic * impulse * j_wc;
    ad -= id_ * impulse * j_wd;
    V2 cs[4] = {ca, cb, cc, cd};
    double as[4] = {aa, ab, ac, ad};
    for (int k = 0; k < 4; ++k) { pv[g.idx[k]].c = cs[k]; pv[g.idx[k]].a = as[k]; }
  }
  return 1;
}

/* ---- wheel (joints/wheel.py:241-417) ---- */
static void wheel_init(World* w, Joint* J, PV* pv, const Step* st) {
  JB b = joint_bodies(w, J);
  double ma = b.ma, mb = b.mb, ia = b.Ia, ib = b.Ib;
  V2 ca = pv[b.ia].c, cb = pv[b.ib].c;
  double aa = pv[b.ia].a, ab = pv[b.ib].a;
  V2 va = pv[b.ia].v, vb = pv[b.ib].v;
  double wa = pv[b.ia].w, wb = pv[b.ib].w;
  M22 qa = rot(aa), qb = rot(ab);
  V2 ra = mulmv(qa, sub(jget2(J, B2G_JD_LA_X), b.lca));
  V2 rb = mulmv(qb, sub(jget2(J, B2G_JD_LB_X), b.lcb));
  jset2(J, B2G_JD_RA_X, ra); jset2(J, B2G_JD_RB_X, rb);
  V2 d = sub(sub(add(cb, rb), ca), ra);
  V2 ay = mulmv(qa, jget2(J, B2G_JD_PERP_X));
  jset2(J, B2G_JD_WHE_AY_X, ay);
  double s_ay = cross(add(d, ra), ay), s_by = cross(rb, ay);
  J->d[B2G_JD_WHE_S_AY] = s_ay; J->d[B2G_JD_WHE_S_BY] = s_by;
  double mass = ma + mb + ia * pysq(s_ay) + ib * pysq(s_by);
  if (mass > 0.0) mass = 1.0 / mass;
  J->d[B2G_JD_WHE_MASS] = mass;
  J->d[B2G_JD_WHE_SPRING_MASS] = 0.0;
  J->d[B2G_JD_WHE_BIAS] = 0.0;
  J->d[B2G_JD_WHE_GAMMA] = 0.0;
  if (J->d[B2G_JD_FREQ] > 0.0) {
    V2 ax = mulmv(qa, jget2(J, B2G_JD_AXIS_X));
    jset2(J, B2G_JD_WHE_AX_X, ax);
    double s_ax = cross(add(d, ra), ax), s_bx = cross(rb, ax);
    J->d[B2G_JD_WHE_S_AX] = s_ax; J->d[B2G_JD_WHE_S_BX] = s_bx;
    double inv_mass = ma + mb + ia * pysq(s_ax) + ib * pysq(s_bx);
    if (inv_mass > 0.0) {
      double spring_mass = 1.0 / inv_mass;
      double C = dot(d, ax);
      double omega = 2.0 * PI_ * J->d[B2G_JD_FREQ];
      double dd = 2.0 * spring_mass * J->d[B2G_JD_DAMP] * omega;
      double k = spring_mass * pysq(omega);
      double dt = st->dt;
      double gamma = dt * (dd + dt * k);
      if (gamma > 0.0) gamma = 1.0 / gamma;
      J->d[B2G_JD_WHE_GAMMA] = gamma;
      J->d[B2G_JD_WHE_BIAS] = C * dt * k * gamma;
      spring_mass = inv_mass + gamma;
      if (spring_mass > 0.0) spring_mass = 1.0 / spring_mass;
      J->d[B2G_JD_WHE_SPRING_MASS] = spring_mass;
    }
  } else {
    J->d[B2G_JD_SPRING_IMP] = 0.0;
  }
  if (JFLAG(J, B2G_JF_MOTOR)) {
    double motor_mass = ia + ib;
    if (motor_mass > 0.0) motor_mass = 1.0 / motor_mass;
    J->d[B2G_JD_WHE_MOTOR_MASS] = motor_mass;
  } else {
    J->d[B2G_JD_WHE_MOTOR_MASS] = 0.0;
    J->d[B2G_JD_MOTOR_IMP] = 0.0;
  }
  if (st->warm_starting) {
    double r = st->dt_ratio;
    double impulse = J->d[B2G_JD_IMP_X] * r;
    double spring_impulse = J->d[B2G_JD_SPRING_IMP] * r;
    double motor_impulse = J->d[B2G_JD_MOTOR_IMP] * r;
    J->d[B2G_JD_IMP_X] = impulse;
    J->d[B2G_JD_SPRING_IMP] = spring_impulse;
    J->d[B2G_JD_MOTOR_IMP] = motor_impulse;
    V2 P = add(scale(impulse, ay), scale(spring_impulse, jget2(J, B2G_JD_WHE_AX_X)));
    double La = impulse * s_ay + spring_impulse * J->d[B2G_JD_WHE_S_AX] + motor_impulse;
    double Lb = impulse * s_by + spring_impulse * J->d[B2G_JD_WHE_S_BX] + motor_impulse;
    va = sub(va, scale(ma, P));
    wa -= ia * La;
    vb = add(vb, scale(mb, P));
    wb += ib * Lb;
  } else {
    J->d[B2G_JD_IMP_X] = 0.0;
    J->d[B2G_JD_SPRING_IMP] = 0.0;
    J->d[B2G_JD_MOTOR_IMP] = 0.0;
  }
  pv[b.ia].v = va; pv[b.ia].w = wa;
  pv[b.ib].v = vb; pv[b.ib].w = wb;
}

static void wheel_solve_velocity(World* w, Joint* J, PV* pv, const Step* st) {
  JB b = joint_bodies(w, J);
  V2 va = pv[b.ia].v, vb = pv[b.ib].v;
  double wa = pv[b.ia].w, wb = pv[b.ib].w;
  double ma = b.ma, mb = b.mb, ia = b.Ia, ib = b.Ib;
  V2 ax = jget2(J, B2G_JD_WHE_AX_X), ay = jget2(J, B2G_JD_WHE_AY_X);
  double s_ax = J->d[B2G_JD_WHE_S_AX], s_bx = J->d[B2G_JD_WHE_S_BX];
  double s_ay = J->d[B2G_JD_WHE_S_AY], s_by = J->d[B2G_JD_WHE_S_BY];
  double Cdot = dot(ax, sub(vb, va)) + s_bx * wb - s_ax * wa;
  double impulse = -J->d[B2G_JD_WHE_SPRING_MASS] *
                   (Cdot + J->d[B2G_JD_WHE_BIAS] + J->d[B2G_JD_WHE_GAMMA] * J->d[B2G_JD_SPRING_IMP]);
  J->d[B2G_JD_SPRING_IMP] += impulse;
  V2 P = scale(impulse, ax);
  double La = impulse * s_ax, Lb = impulse * s_bx;
  va = sub(va, scale(ma, P));
  wa -= ia * La;
  vb = add(vb, scale(mb, P));
  wb += ib * Lb;
  Cdot = wb - wa - J->d[B2G_JD_MOTOR_SPEED];
  impulse = -J->d[B2G_JD_WHE_MOTOR_MASS] * Cdot;
  double old = J->d[B2G_JD_MOTOR_IMP];
  double max_impulse = st->dt * J->d[B2G_JD_MAX_MOTOR];
  J->d[B2G_JD_MOTOR_IMP] = pclamp(old + impulse, -max_impulse, max_impulse);
  impulse = J->d[B2G_JD_MOTOR_IMP] - old;
  wa -= ia * impulse;
  wb += ib * impulse;
  Cdot = dot(ay, sub(vb, va)) + s_by * wb - s_ay * wa;
  impulse = -J->d[B2G_JD_WHE_MASS] * Cdot;
  J->d[B2G_JD_IMP_X] += impulse;
  P = scale(impulse, ay);
  La = impulse * s_ay;
  Lb = impulse * s_by;
  va = sub(va, scale(ma, P));
  wa -= ia * La;
  vb = add(vb, scale(mb, P));
  wb += ib * Lb;
  pv[b.ia].v = va; pv[b.ia].w = wa;
  pv[b.ib].v = vb; pv[b.ib].w = wb;
}

static int wheel_solve_position(World* w, Joint* J, PV* pv) {
  JB b = joint_bodies(w, J);
  V2 ca = pv[b.ia].c, cb = pv[b.ib].c;
  double aa = pv[b.ia].a, ab = pv[b.ib].a;
  double ma = b.ma, mb = b.mb, ia = b.Ia, ib = b.Ib;
  M22 qa = rot(aa), qb = rot(ab);
  V2 ra = mulmv(qa, sub(jget2(J, B2G_JD_LA_X), b.lca));
  V2 rb = mulmv(qb, sub(jget2(J, B2G_JD_LB_X), b.lcb));
  V2 d = sub(add(sub(cb, ca), rb), ra);
  /* the effective mass from the setup's s_ay, s_by (wheel.py:402) */
  double k = ma + mb + ia * pysq(J->d[B2G_JD_WHE_S_AY]) + ib * pysq(J->d[B2G_JD_WHE_S_BY]);
  V2 ay = mulmv(qa, jget2(J, B2G_JD_PERP_X));
  double s_ay = cross(add(d, ra), ay), s_by = cross(rb, ay);
  double C = dot(d, ay);
  double impulse = k != 0.0 ? -C / k : 0.0;
  V2 P = scale(impulse, ay);
  double La = impulse * s_ay, Lb = impulse * s_by;
  ca = sub(ca, scale(ma, P));
  aa -= ia * La;
  cb = add(cb, scale(mb, P));
  ab += ib * Lb;
  pv[b.ia].c = ca; pv[b.ia].a = aa;
  pv[b.ib].c = cb; pv[b.ib].a = ab;
  return fabs(C) <= LINEAR_SLOP;
}

/* ---- weld (joints/weld.py:127-257); the setup's mass matrix has col2.z = c3x (weld.py:162) ---- */
static void weld_init(World* w, Joint* J, PV* pv, const Step* st) {
  JB b = joint_bodies(w, J);
  double ma = b.ma, mb = b.mb, ia = b.Ia, ib = b.Ib;
  double aa = pv[b.ia].a, ab = pv[b.ib].a;
  V2 va = pv[b.ia].v, vb = pv[b.ib].v;
  double wa = pv[b.ia].w, wb = pv[b.ib].w;
  M22 qa = rot(aa), qb = rot(ab);
  V2 ra = mulmv(qa, sub(jget2(J, B2G_JD_LA_X), b.lca));
  V2 rb = mulmv(qb, sub(jget2(J, B2G_JD_LB_X), b.lcb));
  jset2(J, B2G_JD_RA_X, ra); jset2(J, B2G_JD_RB_X, rb);
  double c2x = -ra.y * ra.x * ia - rb.y * rb.x * ib;
  double c3x = -ra.y * ia - rb.y * ib;
  M33 m;
  m.c1 = v3(ma + mb + ra.y * ra.y * ia + rb.y * rb.y * ib, c2x, c3x);
  m.c2 = v3(c2x, ma + mb + ra.x * ra.x * ia + rb.x * rb.x * ib, c3x);
  m.c3 = v3(c3x, ra.x * ia + rb.x * ib, ia + ib);
  jset33(J, B2G_JD_WEL_MASS, m);
  if (st->warm_starting) {
    V3 imp = scale3(st->dt_ratio, jget3(J, B2G_JD_IMP_X));
    jset3(J, B2G_JD_IMP_X, imp);
    V2 P = v2(imp.x, imp.y);
    va = sub(va, scale(ma, P));
    wa -= ia * (cross(ra, P) + imp.z);
    vb = add(vb, scale(mb, P));
    wb += ib * (cross(rb, P) + imp.z);
  } else {
    jset3(J, B2G_JD_IMP_X, v3(0.0, 0.0, 0.0));
  }
  pv[b.ia].v = va; pv[b.ia].w = wa;
  pv[b.ib].v = vb; pv[b.ib].w = wb;
}

static void weld_solve_velocity(World* w, Joint* J, PV* pv) {
  JB b = joint_bodies(w, J);
  V2 va = pv[b.ia].v, vb = pv[b.ib].v;
  double wa = pv[b.ia].w, wb = pv[b.ib].w;
  V2 ra = jget2(J, B2G_JD_RA_X), rb = jget2(J, B2G_JD_RB_X);
  V2 Cdot1 = sub(sub(add(vb, cross_sv(wb, rb)), va), cross_sv(wa, ra));
  double Cdot2 = wb - wa;
  V3 impulse = solve33(jget33(J, B2G_JD_WEL_MASS), neg3(v3(Cdot1.x, Cdot1.y, Cdot2)));
  jset3(J, B2G_JD_IMP_X, add3(jget3(J, B2G_JD_IMP_X), impulse));
  V2 P = v2(impulse.x, impulse.y);
  va = sub(va, scale(b.ma, P));
  wa -= b.Ia * (cross(ra, P) + impulse.z);
  vb = add(vb, scale(b.mb, P));
  wb += b.Ib * (cross(rb, P) + impulse.z);
  pv[b.ia].v = va; pv[b.ia].w = wa;
  pv[b.ib].v = vb; pv[b.ib].w = wb;
}

static int weld_solve_position(World* w, Joint* J, PV* pv) {
  JB b = joint_bodies(w, J);
  V2 ca = pv[b.ia].c, cb = pv[b.ib].c;
  double aa = pv[b.ia].a, ab = pv[b.ib].a;
  double ma = b.ma, mb = b.mb, ia = b.Ia, ib = b.Ib;
  M22 qa = rot(aa), qb = rot(ab);
  V2 ra = mulmv(qa, sub(jget2(J, B2G_JD_LA_X), b.lca));
  V2 rb = mulmv(qb, sub(jget2(J, B2G_JD_LB_X), b.lcb));
  V2 C1 = sub(sub(add(cb, rb), ca), ra);
  double C2 = ab - aa - J->d[B2G_JD_REF];
  double position_error = length_(C1), angular_error = fabs(C2);
  double c2x = -ra.y * ra.x * ia - rb.y * rb.x * ib;
  double c3x = -ra.y * ia - rb.y * ib;
  double c3y = ra.x * ia + rb.x * ib;
  M33 m;
  m.c1 = v3(ma + mb + ra.y * ra.y * ia + rb.y * rb.y * ib, c2x, c3x);
  m.c2 = v3(c2x, ma + mb + ra.x * ra.x * ia + rb.x * rb.x * ib, c3y);
  m.c3 = v3(-ra.y * ia - rb.y * ib, ra.x * ia + rb.x * ib, ia + ib);
  jset33(J, B2G_JD_WEL_MASS, m); /* the joint's _mass is rewritten here (weld.py:230-242) */
  V3 impulse = neg3(solve33(m, v3(C1.x, C1.y, C2)));
  V2 P = v2(impulse.x, impulse.y);
  ca = sub(ca, scale(ma, P));
  aa -= ia * (cross(ra, P) + impulse.z);
  cb = add(cb, scale(mb, P));
  ab += ib * (cross(rb, P) + impulse.z);
  pv[b.ia].c = ca; pv[b.ia].a = aa;
  pv[b.ib].c = cb; pv[b.ib].a = ab;
  return position_error <= LINEAR_SLOP && angular_error <= ANGULAR_SLOP;
}

/* ---- friction (joints/friction.py:115-233); the velocity solve stores its results into the
   position list (friction.py:228-229), so only its impulses (through the warm start) act ---- */
static void friction_init(World* w, Joint* J, PV* pv, const Step* st) {
  JB b = joint_bodies(w, J);
  double ma = b.ma, mb = b.mb, ia = b.Ia, ib = b.Ib;
  double aa = pv[b.ia].a, ab = pv[b.ib].a;
  V2 va = pv[b.ia].v, vb = pv[b.ib].v;
  double wa = pv[b.ia].w, wb = pv[b.ib].w;
  M22 qa = rot(aa), qb = rot(ab);
  V2 ra = mulmv(qa, sub(jget2(J, B2G_JD_LA_X), b.lca));
  V2 rb = mulmv(qb, sub(jget2(J, B2G_JD_LB_X), b.lcb));
  jset2(J, B2G_JD_RA_X, ra); jset2(J, B2G_JD_RB_X, rb);
  M22 K;
  K.c1x = ma + mb + ia * ra.y * ra.y + ib * rb.y * rb.y;
  K.c1y = -ia * ra.x * ra.y - ib * rb.x * rb.y;
  K.c2x = K.c1y;
  K.c2y = ma + mb + ia * ra.x * ra.x + ib * rb.x * rb.x;
  jset22(J, B2G_JD_FRI_LMASS, m22_inverse(K));
  double angular_mass = ia + ib;
  if (angular_mass > 0.0) angular_mass = 1.0 / angular_mass;
  J->d[B2G_JD_FRI_AMASS] = angular_mass;
  if (st->warm_starting) {
    V2 li = scale(st->dt_ratio, jget2(J, B2G_JD_IMP_X));
    double ang = J->d[B2G_JD_IMP_Z] * st->dt_ratio;
    jset2(J, B2G_JD_IMP_X, li);
    J->d[B2G_JD_IMP_Z] = ang;
    va = sub(va, scale(ma, li));
    wa -= ia * (cross(ra, li) + ang);
    vb = add(vb, scale(mb, li));
    wb += ib * (cross(rb, li) + ang);
  } else {
    jset2(J, B2G_JD_IMP_X, v2(0.0, 0.0));
    J->d[B2G_JD_IMP_Z] = 0.0;
  }
  pv[b.ia].v = va; pv[b.ia].w = wa;
  pv[b.ib].v = vb; pv[b.ib].w = wb;
}

static void friction_solve_velocity(World* w, Joint* J, PV* pv, const Step* st) {
  JB b = joint_bodies(w, J);
  V2 va = pv[b.ia].v, vb = pv[b.ib].v;
  double wa = pv[b.ia].w, wb = pv[b.ib].w;
  V2 ra = jget2(J, B2G_JD_RA_X), rb = jget2(J, B2G_JD_RB_X);
  double dt = st->dt;
  double Cdot = wb - wa;
  double impulse = -J->d[B2G_JD_FRI_AMASS] * Cdot;
  double old = J->d[B2G_JD_IMP_Z];
  double max_impulse = dt * J->d[B2G_JD_MAX_TORQUE];
  J->d[B2G_JD_IMP_Z] = pclamp(old + impulse, -max_impulse, max_impulse);
  impulse = J->d[B2G_JD_IMP_Z] - old;
  wa -= b.Ia * impulse;
  wb += b.Ib * impulse;
  V2 Cdot2 = sub(sub(add(vb, cross_sv(wb, rb)), va), cross_sv(wa, ra));
  V2 limp = neg(mulmv(jget22(J, B2G_JD_FRI_LMASS), Cdot2));
  V2 lold = jget2(J, B2G_JD_IMP_X);
  V2 li = add(lold, limp);
  double lmax = dt * J->d[B2G_JD_MAX_MOTOR];
  if (length_sq(li) > pysq(lmax)) li = scale(lmax, normalized(li));
  jset2(J, B2G_JD_IMP_X, li);
  (void)pv; /* the updated velocities are dropped (friction.py:228-229) */
}

/* ---- rope (joints/rope.py:104-248); the attribute _u keeps the unnormalised separation
   (`u /= length` rebinds the local u), and the warm start and velocity solve use it ---- */
static void rope_init(World* w, Joint* J, PV* pv, const Step* st) {
  JB b = joint_bodies(w, J);
  double ma = b.ma, mb = b.mb, ia = b.Ia, ib = b.Ib;
  V2 ca = pv[b.ia].c, cb = pv[b.ib].c;
  double aa = pv[b.ia].a, ab = pv[b.ib].a;
  V2 va = pv[b.ia].v, vb = pv[b.ib].v;
  double wa = pv[b.ia].w, wb = pv[b.ib].w;
  M22 qa = rot(aa), qb = rot(ab);
  V2 ra = mulmv(qa, sub(jget2(J, B2G_JD_LA_X), b.lca));
  V2 rb = mulmv(qb, sub(jget2(J, B2G_JD_LB_X), b.lcb));
  jset2(J, B2G_JD_RA_X, ra); jset2(J, B2G_JD_RB_X, rb);
  V2 u = sub(sub(add(cb, rb), ca), ra);
  jset2(J, B2G_JD_ROP_U_X, u);
  double length = length_(u);
  J->d[B2G_JD_ROP_LENGTH] = length;
  double C = length - J->d[B2G_JD_LENGTH];
  J->i[B2G_JI_STATE] = C > 0.0 ? B2G_LIMIT_AT_UPPER : B2G_LIMIT_INACTIVE;
  if (length > LINEAR_SLOP) {
    u = vdiv(u, length);
  } else {
    J->d[B2G_JD_ROP_MASS] = 0.0;
    J->d[B2G_JD_IMP_X] = 0.0;
    return;
  }
  double cra = cross(ra, u), crb = cross(rb, u);
  double inv_mass = ma + ia * cra * cra + mb + ib * crb * crb;
  J->d[B2G_JD_ROP_MASS] = inv_mass != 0.0 ? 1.0 / inv_mass : 0.0;
  if (st->warm_starting) {
    double imp = J->d[B2G_JD_IMP_X] * st->dt_ratio;
    J->d[B2G_JD_IMP_X] = imp;
    V2 P = scale(imp, jget2(J, B2G_JD_ROP_U_X));
    va = sub(va, scale(ma, P));
    wa -= ia * cross(ra, P);
    vb = add(vb, scale(mb, P));
    wb += ib * cross(rb, P);
  } else {
    J->d[B2G_JD_IMP_X] = 0.0;
  }
  pv[b.ia].v = va; pv[b.ia].w = wa;
  pv[b.ib].v = vb; pv[b.ib].w = wb;
}

static void rope_solve_velocity(World* w, Joint* J, PV* pv, const Step* st) {
  JB b = joint_bodies(w, J);
  V2 va = pv[b.ia].v, vb = pv[b.ib].v;
  double wa = pv[b.ia].w, wb = pv[b.ib].w;
  V2 ra = jget2(J, B2G_JD_RA_X), rb = jget2(J, B2G_JD_RB_X);
  V2 u = jget2(J, B2G_JD_ROP_U_X);
  V2 vpa = add(va, cross_sv(wa, ra)), vpb = add(vb, cross_sv(wb, rb));
  double C = J->d[B2G_JD_ROP_LENGTH] - J->d[B2G_JD_LENGTH];
  double Cdot = dot(u, sub(vpb, vpa));
  if (C < 0.0) Cdot += st->inv_dt * C;
  double impulse = -J->d[B2G_JD_ROP_MASS] * Cdot;
  double old = J->d[B2G_JD_IMP_X];
  J->d[B2G_JD_IMP_X] = pmin(0.0, old + impulse);
  impulse = J->d[B2G_JD_IMP_X] - old;
  V2 P = scale(impulse, u);
  va = sub(va, scale(b.ma, P));
  wa -= b.Ia * cross(ra, P);
  vb = add(vb, scale(b.mb, P));
  wb += b.Ib * cross(rb, P);
  pv[b.ia].v = va; pv[b.ia].w = wa;
  pv[b.ib].v = vb; pv[b.ib].w = wb;
}

static int rope_solve_position(World* w, Joint* J, PV* pv) {
  JB b = joint_bodies(w, J);
  V2 ca = pv[b.ia].c, cb = pv[b.ib].c;
  double aa = pv[b.ia].a, ab = pv[b.ib].a;
  M22 qa = rot(aa), qb = rot(ab);
  V2 ra = mulmv(qa, sub(jget2(J, B2G_JD_LA_X), b.lca));
  V2 rb = mulmv(qb, sub(jget2(J, B2G_JD_LB_X), b.lcb));
  V2 u = sub(sub(add(cb, rb), ca), ra);
  double length = length_(u);
  u = normalized(u);
  double C = length - J->d[B2G_JD_LENGTH];
  C = pclamp(C, 0.0, MAX_LINEAR_CORRECTION);
  double impulse = -J->d[B2G_JD_ROP_MASS] * C;
  V2 P = scale(impulse, u);
  ca = sub(ca, scale(b.ma, P));
  aa -= b.Ia * cross(ra, P);
  cb = add(cb, scale(b.mb, P));
  ab += b.Ib * cross(rb, P);
  pv[b.ia].c = ca; pv[b.ia].a = aa;
  pv[b.ib].c = cb; pv[b.ib].a = ab;
  return length - J->d[B2G_JD_LENGTH] < LINEAR_SLOP;
}

/* Joint._init_velocity_constraints / _solve_velocity_constraints / _solve_position_constraints */
static void joint_init(World* w, Joint* J, PV* pv, const Step* st) {
  switch (J->i[B2G_JI_TYPE]) {
    case B2G_JT_REVOLUTE: revolute_init(w, J, pv, st); break;
    case B2G_JT_PRISMATIC: prismatic_init(w, J, pv, st); break;
    case B2G_JT_DISTANCE: distance_init(w, J, pv, st); break;
    case B2G_JT_PULLEY: pulley_init(w, J, pv, st); break;
    case B2G_JT_MOUSE: mouse_init(w, J, pv, st); break;
    case B2G_JT_GEAR: gear_init(w, J, pv, st); break;
    case B2G_JT_WHEEL: wheel_init(w, J, pv, st); break;
    case B2G_JT_WELD: weld_init(w, J, pv, st); break;
    case B2G_JT_FRICTION: friction_init(w, J, pv, st); break;
    case B2G_JT_ROPE: rope_init(w, J, pv, st); break;
    default: w->err |= ORC_E_UNSUPPORTED;
  }
}
static void joint_solve_velocity(World* w, Joint* J, PV* pv, const Step* st) {
  switch (J->i[B2G_JI_TYPE]) {
    case B2G_JT_REVOLUTE: revolute_solve_velocity(w, J, pv, st); break;
    case B2G_JT_PRISMATIC: prismatic_solve_velocity(w, J, pv, st); break;
    case B2G_JT_DISTANCE: distance_solve_velocity(w, J, pv); break;
    case B2G_JT_PULLEY: pulley_solve_velocity(w, J, pv); break;
    case B2G_JT_MOUSE: mouse_solve_velocity(w, J, pv, st); break;
    case B2G_JT_GEAR: gear_solve_velocity(w, J, pv); break;
    case B2G_JT_WHEEL: wheel_solve_velocity(w, J, pv, st); break;
    case B2G_JT_WELD: weld_solve_velocity(w, J, pv); break;
    case B2G_JT_FRICTION: friction_solve_velocity(w, J, pv, st); break;
    case B2G_JT_ROPE: rope_solve_velocity(w, J, pv, st); break;
    default: break;
  }
}
static int joint_solve_position(World* w, Joint* J, PV* pv) {
  switch (J->i[B2G_JI_TYPE]) {
    case B2G_JT_REVOLUTE: return revolute_solve_position(w, J, pv);
    case B2G_JT_PRISMATIC: return prismatic_solve_position(w, J, pv);
    case B2G_JT_DISTANCE: return distance_solve_position(w, J, pv);
    case B2G_JT_PULLEY: return pulley_solve_position(w, J, pv);
    case B2G_JT_GEAR: return gear_solve_position(w, J, pv);
    case B2G_JT_WHEEL: return wheel_solve_position(w, J, pv);
    case B2G_JT_WELD: return weld_solve_position(w, J, pv);
    case B2G_JT_ROPE: return rope_solve_position(w, J, pv);
    default: return 1; /* mouse, friction (mouse.py:228-229, friction.py:231-233) */
  }
}

static void island_solve(World* w, IslandBuf* is, const Step* st) { /* island.py:65-179 */
  double dt = st->dt;
  PV* pv = is->pv;
  for (int i = 0; i < is->nbodies; ++i) {
    Body* b = &w->b[is->bodies[i]];
    V2 c = b->c;
    double a = b->a;
    V2 v = b->v;
    double wv = b->w;
    b->c0 = c;
    b->a0 = a;
    if (b->type == B2G_DYNAMIC) {
      v = add(v, scale(dt, add(scale(b->gscale, w->gravity), scale(b->inv_mass, b->force))));
      wv += dt * b->invI * b->torque;
      v = scale(pclamp(1.0 - dt * b->ldamp, 0.0, 1.0), v);
      wv *= pclamp(1.0 - dt * b->adamp, 0.0, 1.0);
    }
    pv[i].c = c; pv[i].a = a; pv[i].v = v; pv[i].w = wv;
  }
  int nvc = is->ncontacts;
  cs_build(w, is, st);
  cs_init_velocity(w, is);
  /* warm_start (contact.py:498-525) */
  if (st->warm_starting) {
    for (int k = 0; k < nvc; ++k) {
      VC* vc = &is->vc[k];
      V2 normal = vc->normal, tangent = cross_vs(normal, 1.0);
      V2 va = pv[vc->ia].v, vb = pv[vc->ib].v;
      double wa = pv[vc->ia].w, wb = pv[vc->ib].w;
      for (int j = 0; j < vc->npts; ++j) {
        VCP* p = &vc->p[j];
        V2 P = add(scale(p->ni, normal), scale(p->ti, tangent));
        wa -= vc->Ia * cross(p->ra, P);
        va = sub(va, scale(vc->ma, P));
        wb += vc->Ib * cross(p->rb, P);
        vb = add(vb, scale(vc->mb, P));
      }
      pv[vc->ia].v = va; pv[vc->ia].w = wa;
      pv[vc->ib].v = vb; pv[vc->ib].w = wb;
    }
  }
  for (int k = 0; k < is->njoints; ++k) joint_init(w, &w->j[is->joints[k]], pv, st); /* island.py:107-108 */
  for (int it = 0; it < st->vel_iters; ++it) { /* island.py:111-114: joints, then contacts */
    for (int k = 0; k < is->njoints; ++k) joint_solve_velocity(w, &w->j[is->joints[k]], pv, st);
    solve_velocity(is->vc, nvc, pv);
  }
  /* store_impulses (contact.py:801-806): zip(manifold.used_points, vc.points) */
  for (int k = 0; k < nvc; ++k) {
    VC* vc = &is->vc[k];
    Contact* c = &w->c[vc->contact];
    int n = c->m.pc < vc->npts ? c->m.pc : vc->npts;
    for (int j = 0; j < n; ++j) { c->m.pts[j].ni = vc->p[j].ni; c->m.pts[j].ti = vc->p[j].ti; }
  }
  cs_report(w, is, 2, w->post_key);
  w->post_key += nvc;
  /* integrate positions (island.py:120-138) */
  for (int i = 0; i < is->nbodies; ++i) {
    V2 v = pv[i].v;
    double wv = pv[i].w;
    V2 tr = scale(dt, v);
    if (dot(tr, tr) > MAX_TRANSLATION_SQR) {
      double ratio = MAX_TRANSLATION / length_(tr);
      v = scale(ratio, v);
    }
    double rotation = dt * wv;
    if (pysq(rotation) > MAX_ROTATION_SQR) {
      double ratio = MAX_ROTATION / fabs(rotation);
      wv *= ratio;
    }
    pv[i].c = add(pv[i].c, scale(dt, v));
    pv[i].a += dt * wv;
    pv[i].v = v;
    pv[i].w = wv;
  }
  int position_solved = 0;
  for (int it = 0; it < st->pos_iters; ++it) { /* island.py:141-152: contacts, then every joint */
    int contacts_okay = solve_position(is->pc, nvc, pv);
    int joints_okay = 1;
    for (int k = 0; k < is->njoints; ++k) {
      int ok = joint_solve_position(w, &w->j[is->joints[k]], pv);
      joints_okay = joints_okay && ok;
    }
    if (contacts_okay && joints_okay) { position_solved = 1; break; }
  }
  for (int i = 0; i < is->nbodies; ++i) {
    Body* b = &w->b[is->bodies[i]];
    b->c = pv[i].c; b->a = pv[i].a; b->v = pv[i].v; b->w = pv[i].w;
    b->xf = xf_from(b->c, b->a, b->lc); /* _synchronize_transform */
  }
  if (w->allow_sleep) { /* island.py:161-179, quirk Q7 */
    double min_sleep = MAX_FLOAT;
    const Body* stale = &w->b[is->bodies[is->nbodies - 1]];
    int last_ns = -1;
    for (int i = 0; i < is->nbodies; ++i) {
      Body* b = &w->b[is->bodies[i]];
      if (b->type == B2G_STATIC) continue;
      last_ns = is->bodies[i];
      if (!stale->allow_sleep || pysq(b->w) > ANGULAR_SLEEP_TOLERANCE_SQR ||
          length_sq(b->v) > LINEAR_SLEEP_TOLERANCE_SQR) {
        b->sleep_time = 0.0;
        min_sleep = 0.0;
      } else {
        b->sleep_time += dt;
        min_sleep = pmin(min_sleep, b->sleep_time);
      }
    }
    if (min_sleep >= TIME_TO_SLEEP && position_solved && last_ns >= 0) set_awake(&w->b[last_ns], 0);
  }
}

/* Body._synchronize_fixtures (body.py:898-904): each fixture's swept AABB from the sweep's
   start (c0, a0) to the current transform; a proxy that leaves its fat AABB is re-fattened and
   appended to the move buffer (Fixture._synchronize, fixture.py:287-298; DynamicTree.move_proxy,
   dynamictree.py:152-188).  Returns nonzero on allocation failure. */
static int sync_body_fixtures(World* w, int bi) {
  Body* b = &w->b[bi];
  XF xf1;
  xf1.R = rot(b->a0);
  xf1.p = sub(b->c0, mulmv(xf1.R, b->lc));
  for (int k = w->bfix_off[bi]; k < w->bfix_off[bi + 1]; ++k) {
    int fi = w->bfix[k];
    Fix* f = &w->f[fi];
    AABB a1 = compute_aabb(&f->sh, xf1);
    AABB a2 = compute_aabb(&f->sh, b->xf);
    AABB aabb = aabb_union(a1, a2);
    V2 disp = sub(b->xf.p, xf1.p);
    if (aabb_contains(&f->fat, &aabb)) continue;
    AABB nb_;
    V2 r = v2(AABB_EXTENSION, AABB_EXTENSION);
    nb_.lo = sub(aabb.lo, r);
    nb_.hi = add(aabb.hi, r);
    V2 d = scale(AABB_MULTIPLIER, disp);
    if (d.x < 0.0) nb_.lo = add(nb_.lo, v2(d.x, 0.0));
    else nb_.hi = add(nb_.hi, v2(d.x, 0.0));
    if (d.y < 0.0) nb_.lo = add(nb_.lo, v2(0.0, d.y));
    else nb_.hi = add(nb_.hi, v2(0.0, d.y));
    f->fat = nb_;
    if (w->nmove == w->mcap) {
      int cap = w->mcap ? 2 * w->mcap : 16;
      int* nm = (int*)realloc(w->move, sizeof(int) * (size_t)cap);
      if (!nm) { w->err |= ORC_E_ALLOC; return -1; }
      w->move = nm;
      w->mcap = cap;
    }
    w->move[w->nmove++] = fi;
  }
  return 0;
}

static void solve_islands(World* w, const Step* st, int then_sync);
static void solve(World* w, const Step* st) { solve_islands(w, st, 1); }

/* world.py:475-592; with then_sync = 0 only the island search and the island solves
   (the solver seam, orc_solve_islands) */
static void solve_islands(World* w, const Step* st, int then_sync) {
  int nb = w->nb, nc = w->nc;
  w->post_key = 0;
  for (int i = 0; i < nb; ++i) { w->b[i].island_flag = 0; w->b[i].island_id = -1; }
  for (int i = 0; i < nc; ++i) w->c[i].island_flag = 0;
  for (int i = 0; i < w->nj; ++i) w->j[i].island_flag = 0;
  /* per-body contact lists in ContactManager order (= body._contacts order) */
  int* off = (int*)calloc((size_t)nb + 1, sizeof(int));
  int* adj = (int*)malloc(sizeof(int) * (size_t)(2 * nc + 1));
  int* fill = (int*)calloc((size_t)nb + 1, sizeof(int));
  IslandBuf is;
  is.bodies = (int*)malloc(sizeof(int) * (size_t)(nb + 1));
  is.contacts = (int*)malloc(sizeof(int) * (size_t)(nc + 1));
  is.joints = (int*)malloc(sizeof(int) * (size_t)(w->nj + 1));
  is.pv = (PV*)malloc(sizeof(PV) * (size_t)(nb + 1));
  is.vc = (VC*)malloc(sizeof(VC) * (size_t)(nc + 1));
  is.pc = (PC*)malloc(sizeof(PC) * (size_t)(nc + 1));
  int* stack = (int*)malloc(sizeof(int) * (size_t)(nb + 1));
  if (!off || !adj || !fill || !is.bodies || !is.contacts || !is.joints || !is.pv || !is.vc || !is.pc || !stack) {
    w->err |= ORC_E_ALLOC;
    goto done;
  }
  for (int i = 0; i < nc; ++i) {
    off[w->f[w->c[i].fa].body + 1]++;
    off[w->f[w->c[i].fb].body + 1]++;
  }
  for (int i = 0; i < nb; ++i) off[i + 1] += off[i];
  for (int i = 0; i < nc; ++i) {
    int ba = w->f[w->c[i].fa].body, bb = w->f[w->c[i].fb].body;
    adj[off[ba] + fill[ba]++] = i;
    adj[off[bb] + fill[bb]++] = i;
  }
  int island_id = 0;
  for (int seed = 0; seed < nb; ++seed) {
    Body* sb = &w->b[seed];
    if (sb->island_flag) continue;
    if (!sb->awake || !sb->active) continue;
    if (sb->type == B2G_STATIC) continue;
    is.nbodies = 0;
    is.ncontacts = 0;
    is.njoints = 0;
    int sp = 0;
    stack[sp++] = seed;
    sb->island_flag = 1;
    while (sp > 0) {
      int bi = stack[--sp];
      Body* b = &w->b[bi];
      b->island_index = is.nbodies;
      b->island_id = island_id;
      is.bodies[is.nbodies++] = bi;
      set_awake(b, 1);
      if (b->type == B2G_STATIC) continue;
      /* the filtered list is taken before the loop (world.py:528-533) */
      int start = is.ncontacts;
      for (int k = off[bi]; k < off[bi + 1]; ++k) {
        Contact* c = &w->c[adj[k]];
        if (!c->island_flag && c->enabled && c->touching && !w->f[c->fa].sensor && !w->f[c->fb].sensor)
          is.contacts[is.ncontacts++] = adj[k];
      }
      int end = is.ncontacts;
      for (int k = start; k < end; ++k) {
        Contact* c = &w->c[is.contacts[k]];
        c->island_flag = 1;
        int other = other_body(w, c, bi);
        if (w->b[other].island_flag) continue;
        stack[sp++] = other;
        w->b[other].island_flag = 1;
      }
      /* the body's joints (world.py:548-572) */
      for (int k = w->bj_off[bi]; k < w->bj_off[bi + 1]; ++k) {
        Joint* J = &w->j[w->bj[k]];
        if (J->island_flag) continue;
        int other = J->i[B2G_JI_BODY_A] == bi ? J->i[B2G_JI_BODY_B] : J->i[B2G_JI_BODY_A]; /* joint.py:76-81 */
        if (other < 0) { /* a joint on one body (mouse) */
          is.joints[is.njoints++] = w->bj[k];
          J->island_flag = 1;
          continue;
        } else if (!w->b[other].active) {
          continue;
        }
        is.joints[is.njoints++] = w->bj[k];
        J->island_flag = 1;
        if (w->b[other].island_flag) continue;
        stack[sp++] = other;
        w->b[other].island_flag = 1;
      }
    }
    island_solve(w, &is, st);
    for (int i = 0; i < is.nbodies; ++i)
      if (w->b[is.bodies[i]].type == B2G_STATIC) w->b[is.bodies[i]].island_flag = 0;
    ++island_id;
  }
  if (!then_sync) goto done;
  /* Body._synchronize_fixtures for island bodies (world.py:584-589, body.py:898-904) */
  for (int bi = 0; bi < nb; ++bi) {
    Body* b = &w->b[bi];
    if (!b->island_flag || b->type == B2G_STATIC) continue;
    if (sync_body_fixtures(w, bi)) goto done;
  }
  find_new_contacts(w);
done:
  free(off); free(adj); free(fill); free(stack);
  free(is.bodies); free(is.contacts); free(is.joints); free(is.pv); free(is.vc); free(is.pc);
}


/* ------------------------------------------------------------------ */
/* Continuous physics: GJK distance (distance.py), time of impact      */
/* (toi.py), Island.solve_toi (island.py:181-243), World.solve_toi     */
/* (world.py:594-821).                                                 */
/* ------------------------------------------------------------------ */
static const int DISTANCE_MAX_ITERS = 20;  /* settings.py:129 */
static const int TOI_MAX_ITERS = 20;       /* settings.py:82 */
static const int MAX_SUB_STEPS = 8;        /* settings.py:73 */
static const int MAX_TOI_CONTACTS = 32;    /* settings.py:79 */
static const double TOI_BAUMGARTE = 0.75;  /* settings.py:110 */
/* The reference's inner TOI loop has no bound on its push-back iterations (toi.py:335-339
   leaves the check commented out); the oracle stops at this many and flags the world. */
static const int TOI_PUSH_BACK_LIMIT = 1000;
#define ORC_E_TOI 8

typedef struct { V2 v[B2G_MAX_POLY_VERTS]; int n; double radius; } DProxy;

/* DistanceProxy (distance.py:355-400): circles give their centre, polygons their vertices,
   edges their two vertices */
static void dproxy_init(DProxy* p, const Shape* s) {
  p->radius = s->radius;
  if (s->type == B2G_SHAPE_CIRCLE) { p->v[0] = s->p; p->n = 1; }
  else if (s->type == B2G_SHAPE_POLYGON) { p->n = s->n; for (int i = 0; i < s->n; ++i) p->v[i] = s->v[i]; }
  else { p->v[0] = s->e1; p->v[1] = s->e2; p->n = 2; }
}

/* get_support: the first vertex of largest dot (distance.py:402-405) */
static int dproxy_support(const DProxy* p, V2 d) {
  int best = 0;
  double bv = dot(p->v[0], d);
  for (int i = 1; i < p->n; ++i) {
    double x = dot(p->v[i], d);
    if (x > bv) { bv = x; best = i; }
  }
  return best;
}

typedef struct { V2 wa, wb, w; double a; int ia, ib; } SVert;
typedef struct { SVert v[3]; int count; } Simplex;
typedef struct { double metric; int count; int ia[3], ib[3]; } SCache;

static double simplex_metric(const Simplex* s) { /* distance.py:180-193 */
  if (s->count == 1) return 0.0;
  if (s->count == 2) return length_(sub(s->v[0].w, s->v[1].w));
  return cross(sub(s->v[1].w, s->v[0].w), sub(s->v[2].w, s->v[0].w));
}

static void simplex_vertex(SVert* v, const DProxy* pa, XF xfa, const DProxy* pb, XF xfb) {
  v->wa = xmul(xfa, pa->v[v->ia]);
  v->wb = xmul(xfb, pb->v[v->ib]);
  v->w = sub(v->wb, v->wa);
}

static void simplex_read(Simplex* s, const SCache* c, const DProxy* pa, XF xfa, const DProxy* pb, XF xfb) {
  s->count = c->count; /* distance.py:88-120 */
  for (int i = 0; i < s->count; ++i) {
    SVert* v = &s->v[i];
    v->ia = c->ia[i];
    v->ib = c->ib[i];
    simplex_vertex(v, pa, xfa, pb, xfb);
    v->a = 0.0;
  }
  if (s->count > 1) {
    double m1 = c->metric, m2 = simplex_metric(s);
    if (m2 < 0.5 * m1 || 2.0 * m1 < m2 || m2 < EPSILON) s->count = 0;
  }
  if (s->count == 0) {
    s->v[0].ia = 0;
    s->v[0].ib = 0;
    simplex_vertex(&s->v[0], pa, xfa, pb, xfb);
    s->count = 1;
  }
}

static V2 simplex_closest(const Simplex* s) { /* distance.py:128-138 */
  if (s->count == 1) return s->v[0].w;
  if (s->count == 2) return add(scale(s->v[0].a, s->v[0].w), scale(s->v[1].a, s->v[1].w));
  return v2(0.0, 0.0);
}

static V2 simplex_direction(const Simplex* s) { /* distance.py:140-154 */
  if (s->count == 1) return neg(s->v[0].w);
  V2 e12 = sub(s->v[1].w, s->v[0].w);
  double sgn = cross(e12, neg(s->v[0].w));
  return sgn > 0.0 ? cross_sv(1.0, e12) : cross_vs(e12, 1.0);
}

static void simplex_witness(const Simplex* s, V2* pa, V2* pb) { /* distance.py:156-178 */
  if (s->count == 1) { *pa = s->v[0].wa; *pb = s->v[0].wb; return; }
  if (s->count == 2) {
    *pa = add(scale(s->v[0].a, s->v[0].wa), scale(s->v[1].a, s->v[1].wa));
    *pb = add(scale(s->v[0].a, s->v[0].wb), scale(s->v[1].a, s->v[1].wb));
    return;
  }
  V2 p = add(add(scale(s->v[0].a, s->v[0].wa), scale(s->v[1].a, s->v[1].wa)), scale(s->v[2].a, s->v[2].wa));
  *pa = p;
  *pb = p;
}

static void simplex_solve2(Simplex* s) { /* distance.py:195-245 */
  V2 w1 = s->v[0].w, w2 = s->v[1].w, e12 = sub(w2, w1);
  double d12_2 = -dot(w1, e12);
  if (d12_2 <= 0.0) { s->v[0].a = 1.0; s->count = 1; return; }
  double d12_1 = dot(w2, e12);
  if (d12_1 <= 0.0) { s->v[1].a = 1.0; s->count = 1; s->v[0] = s->v[1]; return; }
  double inv = 1.0 / (d12_1 + d12_2);
  s->v[0].a = d12_1 * inv;
  s->v[1].a = d12_2 * inv;
  s->count = 2;
}

static void simplex_solve3(Simplex* s) { /* distance.py:247-353 */
  V2 w1 = s->v[0].w, w2 = s->v[1].w, w3 = s->v[2].w;
  V2 e12 = sub(w2, w1);
  double d12_1 = dot(w2, e12), d12_2 = -dot(w1, e12);
  V2 e13 = sub(w3, w1);
  double d13_1 = dot(w3, e13), d13_2 = -dot(w1, e13);
  V2 e23 = sub(w3, w2);
  double d23_1 = dot(w3, e23), d23_2 = -dot(w2, e23);
  double n123 = cross(e12, e13);
  double d123_1 = n123 * cross(w2, w3), d123_2 = n123 * cross(w3, w1), d123_3 = n123 * cross(w1, w2);
  if (d12_2 <= 0.0 && d13_2 <= 0.0) { s->v[0].a = 1.0; s->count = 1; return; }
  if (d12_1 > 0.0 && d12_2 > 0.0 && d123_3 <= 0.0) {
    double inv = 1.0 / (d12_1 + d12_2);
    s->v[0].a = d12_1 * inv; s->v[1].a = d12_2 * inv; s->count = 2;
    return;
  }
  if (d13_1 > 0.0 && d13_2 > 0.0 && d123_2 <= 0.0) {
    double inv = 1.0 / (d13_1 + d13_2);
    s->v[0].a = d13_1 * inv; s->v[2].a = d13_2 * inv; s->count = 2;
    s->v[1] = s->v[2];
    return;
  }
  if (d12_1 <= 0.0 && d23_2 <= 0.0) { s->v[1].a = 1.0; s->count = 1; s->v[0] = s->v[1]; return; }
  if (d13_1 <= 0.0 && d23_1 <= 0.0) { s->v[2].a = 1.0; s->count = 1; s->v[0] = s->v[2]; return; }
  if (d23_1 > 0.0 && d23_2 > 0.0 && d123_1 <= 0.0) {
    double inv = 1.0 / (d23_1 + d23_2);
    s->v[1].a = d23_1 * inv; s->v[2].a = d23_2 * inv; s->count = 2;
    s->v[0] = s->v[2];
    return;
  }
  double inv = 1.0 / (d123_1 + d123_2 + d123_3);
  s->v[0].a = d123_1 * inv; s->v[1].a = d123_2 * inv; s->v[2].a = d123_3 * inv;
  s->count = 3;
}

/* shape_distance (distance.py:424-561): GJK; the simplex cache is read and written.
   Returns the distance (without the radii: use_radii False). */
static double gjk_distance(const DProxy* pa, const DProxy* pb, XF xfa, XF xfb, SCache* cache) {
  Simplex s;
  memset(&s, 0, sizeof(s));
  simplex_read(&s, cache, pa, xfa, pb, xfb);
  int iter = 0;
  while (iter < DISTANCE_MAX_ITERS) {
    int save_n = s.count, save_a[3], save_b[3];
    for (int i = 0; i < s.count; ++i) { save_a[i] = s.v[i].ia; save_b[i] = s.v[i].ib; }
    if (s.count == 2) simplex_solve2(&s);
    else if (s.count == 3) { simplex_solve3(&s); break; }
    V2 d = simplex_direction(&s);
    if (length_sq(d) < EPSILON_SQR) break;
    SVert* v = &s.v[s.count];
    v->ia = dproxy_support(pa, mulmtv(xfa.R, neg(d)));
    v->wa = xmul(xfa, pa->v[v->ia]);
    v->ib = dproxy_support(pb, mulmtv(xfb.R, d));
    v->wb = xmul(xfb, pb->v[v->ib]);
    v->w = sub(v->wb, v->wa);
    ++iter;
    int dup = 0;
    for (int i = 0; i < save_n; ++i) if (save_a[i] == v->ia && save_b[i] == v->ib) dup = 1;
    if (dup) break;
    ++s.count;
  }
  V2 wa, wb;
  simplex_witness(&s, &wa, &wb);
  double dist = length_(sub(wa, wb));
  cache->metric = simplex_metric(&s);
  cache->count = s.count;
  for (int i = 0; i < s.count; ++i) { cache->ia[i] = s.v[i].ia; cache->ib[i] = s.v[i].ib; }
  return dist;
}

typedef struct { V2 c0, c, lc; double a0, a, alpha0; } SweepT;

static XF sweep_xf(const SweepT* sw, double beta) { /* Sweep.get_transform (common.py:895-906) */
  V2 p = add(scale(1.0 - beta, sw->c0), scale(beta, sw->c));
  double ang = (1.0 - beta) * sw->a0 + beta * sw->a;
  return xf_from(p, ang, sw->lc);
}

static void sweep_advance(SweepT* sw, double alpha) { /* Sweep.advance (common.py:908-917) */
  double beta = (alpha - sw->alpha0) / (1.0 - sw->alpha0);
  sw->c0 = add(scale(1.0 - beta, sw->c0), scale(beta, sw->c));
  sw->a0 = (1.0 - beta) * sw->a0 + beta * sw->a;
  sw->alpha0 = alpha;
}

static SweepT body_sweep(const Body* b) {
  SweepT s;
  s.c0 = b->c0; s.c = b->c; s.lc = b->lc; s.a0 = b->a0; s.a = b->a; s.alpha0 = b->alpha0;
  return s;
}

static void set_body_sweep(Body* b, const SweepT* s) {
  b->c0 = s->c0; b->c = s->c; b->a0 = s->a0; b->a = s->a; b->alpha0 = s->alpha0;
}

enum { SEP_POINTS, SEP_FACE_A, SEP_FACE_B };
typedef struct { const DProxy *pa, *pb; SweepT sa, sb; int type; V2 axis, lp; } SepFn;

static void sep_init(SepFn* f, const SCache* cache, double t1) { /* toi.py:54-122 */
  XF xfa = sweep_xf(&f->sa, t1), xfb = sweep_xf(&f->sb, t1);
  int ia0 = cache->ia[0], ib0 = cache->ib[0];
  if (cache->count == 1) {
    f->type = SEP_POINTS;
    /* the axis is left unnormalised: `s = self.axis.normalized` discards the unit vector */
    f->axis = sub(xmul(xfb, f->pb->v[ib0]), xmul(xfa, f->pa->v[ia0]));
    return;
  }
  int ia1 = cache->ia[1], ib1 = cache->ib[1];
  if (ia0 == ia1) {
    f->type = SEP_FACE_B;
    V2 b1 = f->pb->v[ib0], b2 = f->pb->v[ib1];
    f->axis = normalized(cross_vs(sub(b2, b1), 1.0));
    V2 n = mulmv(xfb.R, f->axis);
    f->lp = scale(0.5, add(b1, b2));
    V2 pb = xmul(xfb, f->lp), pa = xmul(xfa, f->pa->v[ia0]);
    if (dot(sub(pa, pb), n) < 0.0) f->axis = neg(f->axis);
  } else {
    f->type = SEP_FACE_A;
    V2 a1 = f->pa->v[ia0], a2 = f->pa->v[ia1];
    f->axis = normalized(cross_vs(sub(a2, a1), 1.0));
    V2 n = mulmv(xfa.R, f->axis);
    f->lp = scale(0.5, add(a1, a2));
    V2 pa = xmul(xfa, f->lp), pb = xmul(xfb, f->pb->v[ib0]);
    if (dot(sub(pb, pa), n) < 0.0) f->axis = neg(f->axis);
  }
}

/* find_min_separation (toi.py:124-175) and evaluate (177-213): the witness indices are
   found (find) or given (evaluate) */
static double sep_eval(const SepFn* f, int find, int* ia, int* ib, double t) {
  XF xfa = sweep_xf(&f->sa, t), xfb = sweep_xf(&f->sb, t);
  if (f->type == SEP_POINTS) {
    if (find) {
      *ia = dproxy_support(f->pa, mulmtv(xfa.R, f->axis));
      *ib = dproxy_support(f->pb, mulmtv(xfb.R, neg(f->axis)));
    }
    V2 pa = xmul(xfa, f->pa->v[*ia]), pb = xmul(xfb, f->pb->v[*ib]);
    return dot(sub(pb, pa), f->axis);
  }
  if (f->type == SEP_FACE_A) {
    V2 n = mulmv(xfa.R, f->axis);
    V2 pa = xmul(xfa, f->lp);
    if (find) { *ia = -1; *ib = dproxy_support(f->pb, mulmtv(xfb.R, neg(n))); }
    V2 pb = xmul(xfb, f->pb->v[*ib]);
    return dot(sub(pb, pa), n);
  }
  V2 n = mulmv(xfb.R, f->axis);
  V2 pb = xmul(xfb, f->lp);
  if (find) { *ib = -1; *ia = dproxy_support(f->pa, mulmtv(xfa.R, neg(n))); }
  V2 pa = xmul(xfa, f->pa->v[*ia]);
  return dot(sub(pa, pb), n);
}

enum { TOI_UNKNOWN, TOI_FAILED, TOI_OVERLAPPED, TOI_TOUCHING, TOI_SEPARATED };

/* time_of_impact (toi.py:215-391) over [0, 1]; returns the state, *t_out the time */
static int time_of_impact(World* w, const DProxy* pa, const DProxy* pb, const SweepT* swa, const SweepT* swb,
                          double* t_out) {
  const double t_max = 1.0;
  w->toi_calls++;
  SepFn f;
  f.pa = pa; f.pb = pb;
  f.sa = *swa; f.sb = *swb;
  /* Sweep.normalize (common.py:919-923) on the copies */
  double d = 2.0 * PI_ * floor(f.sa.a0 / (2.0 * PI_));
  f.sa.a0 -= d; f.sa.a -= d;
  d = 2.0 * PI_ * floor(f.sb.a0 / (2.0 * PI_));
  f.sb.a0 -= d; f.sb.a -= d;
  double total_radius = pa->radius + pb->radius;
  double target = pmax(LINEAR_SLOP, total_radius - 3.0 * LINEAR_SLOP);
  double tolerance = 0.25 * LINEAR_SLOP;
  double t1 = 0.0, out_t = t_max;
  int iter = 0, state = TOI_UNKNOWN;
  SCache cache;
  memset(&cache, 0, sizeof(cache));
  for (;;) {
    XF xfa = sweep_xf(&f.sa, t1), xfb = sweep_xf(&f.sb, t1);
    double dist = gjk_distance(pa, pb, xfa, xfb, &cache);
    if (dist <= 0.0) { state = TOI_OVERLAPPED; out_t = 0.0; break; }
    if (dist < target + tolerance) { state = TOI_TOUCHING; out_t = t1; break; }
    sep_init(&f, &cache, t1);
    int done = 0, push_back = 0;
    double t2 = t_max;
    for (;;) {
      int ia = 0, ib = 0;
      double s2 = sep_eval(&f, 1, &ia, &ib, t2);
      if (s2 > target + tolerance) { state = TOI_SEPARATED; out_t = t_max; done = 1; break; }
      if (s2 > target - tolerance) { t1 = t2; break; }
      double s1 = sep_eval(&f, 0, &ia, &ib, t1);
      if (s1 < target - tolerance) { state = TOI_FAILED; out_t = t1; done = 1; break; }
      if (s1 <= target + tolerance) { state = TOI_TOUCHING; out_t = t1; done = 1; break; }
      int rc = 0;
      double a1 = t1, a2 = t2;
      for (;;) {
        double t = (rc & 1) ? a1 + (target - s1) * (a2 - a1) / (s2 - s1) : 0.5 * (a1 + a2);
        double sv = sep_eval(&f, 0, &ia, &ib, t);
        if (fabs(sv - target) < tolerance) { t2 = t; break; }
        if (sv > target) { a1 = t; s1 = sv; }
        else { a2 = t; s2 = sv; }
        ++rc;
        if (rc == 50) break;
      }
      if (++push_back == TOI_PUSH_BACK_LIMIT) { w->err |= ORC_E_TOI; state = TOI_FAILED; out_t = t1; done = 1; break; }
    }
    ++iter;
    if (done) break;
    if (iter == TOI_MAX_ITERS) { state = TOI_FAILED; out_t = t1; break; }
  }
  *t_out = out_t;
  return state;
}

/* Body._advance (body.py:930-937) */
static void body_advance(Body* b, double alpha) {
  SweepT sw = body_sweep(b);
  sweep_advance(&sw, alpha);
  set_body_sweep(b, &sw);
  b->c = b->c0;
  b->a = b->a0;
  b->xf = xf_from(b->c, b->a, b->lc);
}

/* ContactSolver.solve_toi_position_constraints (contact.py:866-933): only the two TOI bodies
   move */
static int solve_toi_position(PC* pcs, int npc, PV* pv, int toi_a, int toi_b) {
  double min_sep = 0.0;
  for (int k = 0; k < npc; ++k) {
    PC* pc = &pcs[k];
    V2 ca = pv[pc->ia].c, cb = pv[pc->ib].c;
    double aa = pv[pc->ia].a, ab = pv[pc->ib].a;
    int ma = pc->ia == toi_a || pc->ia == toi_b, mb = pc->ib == toi_a || pc->ib == toi_b;
    double ima = ma ? pc->ma : 0.0, iIa = ma ? pc->Ia : 0.0;
    double imb = mb ? pc->mb : 0.0, iIb = mb ? pc->Ib : 0.0;
    for (int j = 0; j < pc->npts; ++j) {
      XF xfa = xf_from(ca, aa, pc->lca), xfb = xf_from(cb, ab, pc->lcb);
      V2 normal, point;
      double sep;
      solver_manifold(pc, xfa, xfb, j, &normal, &point, &sep);
      V2 ra = sub(point, ca), rb = sub(point, cb);
      min_sep = pmin(min_sep, sep);
      double C = pclamp(TOI_BAUMGARTE * (sep + LINEAR_SLOP), -MAX_LINEAR_CORRECTION, 0.0);
      double rna = pysq(cross(ra, normal)), rnb = pysq(cross(rb, normal));
      double K = ima + imb + iIa * rna + iIb * rnb;
      double impulse = K > 0.0 ? -C / K : 0.0;
      V2 P = scale(impulse, normal);
      ca = sub(ca, scale(ima, P));
      aa -= iIa * cross(ra, P);
      cb = add(cb, scale(imb, P));
      ab += iIb * cross(rb, P);
    }
    pv[pc->ia].c = ca; pv[pc->ia].a = aa;
    pv[pc->ib].c = cb; pv[pc->ib].a = ab;
  }
  return min_sep >= -1.5 * LINEAR_SLOP;
}

/* Island.solve_toi (island.py:181-243) */
static void island_solve_toi(World* w, IslandBuf* is, const Step* st, int toi_a, int toi_b) {
  PV* pv = is->pv;
  for (int i = 0; i < is->nbodies; ++i) {
    Body* b = &w->b[is->bodies[i]];
    pv[i].c = b->c; pv[i].a = b->a; pv[i].v = b->v; pv[i].w = b->w;
  }
  cs_build(w, is, st);
  for (int it = 0; it < st->pos_iters; ++it)
    if (solve_toi_position(is->pc, is->ncontacts, pv, toi_a, toi_b)) break;
  Body *A = &w->b[is->bodies[toi_a]], *B = &w->b[is->bodies[toi_b]];
  A->c0 = pv[toi_a].c; A->a0 = pv[toi_a].a;
  B->c0 = pv[toi_b].c; B->a0 = pv[toi_b].a;
  cs_init_velocity(w, is);
  for (int it = 0; it < st->vel_iters; ++it) solve_velocity(is->vc, is->ncontacts, pv);
  double dt = st->dt;
  for (int i = 0; i < is->nbodies; ++i) {
    Body* b = &w->b[is->bodies[i]];
    V2 v = pv[i].v;
    double wv = pv[i].w;
    V2 tr = scale(dt, v);
    if (length_sq(tr) > MAX_TRANSLATION_SQR) v = scale(MAX_TRANSLATION / length_(tr), v);
    double rotation = dt * wv;
    if (pysq(rotation) > MAX_ROTATION_SQR) wv *= MAX_ROTATION / fabs(rotation);
    pv[i].c = add(pv[i].c, scale(dt, v));
    pv[i].a += dt * wv;
    b->c = pv[i].c; b->a = pv[i].a; b->v = v; b->w = wv;
    b->xf = xf_from(b->c, b->a, b->lc);
  }
  cs_report(w, is, 3, 0);
}

/* World.solve_toi (world.py:594-821) */
static void solve_toi(World* w, const Step* st) {
  IslandBuf is;
  memset(&is, 0, sizeof(is));
  is.bodies = (int*)malloc(sizeof(int) * 2 * (size_t)MAX_TOI_CONTACTS);
  is.contacts = (int*)malloc(sizeof(int) * (size_t)MAX_TOI_CONTACTS);
  is.pv = (PV*)malloc(sizeof(PV) * 2 * (size_t)MAX_TOI_CONTACTS);
  is.vc = (VC*)malloc(sizeof(VC) * (size_t)MAX_TOI_CONTACTS);
  is.pc = (PC*)malloc(sizeof(PC) * (size_t)MAX_TOI_CONTACTS);
  if (!is.bodies || !is.contacts || !is.pv || !is.vc || !is.pc) { w->err |= ORC_E_ALLOC; goto out; }
  w->toi_mode = 1;
  for (int i = 0; i < w->nb; ++i) { w->b[i].toi_flag = 0; w->b[i].alpha0 = 0.0; }
  for (int i = 0; i < w->nc; ++i) {
    Contact* c = &w->c[i];
    c->toi_flag = 0; c->toi_island = 0; c->toi_count = 0; c->toi = 1.0;
  }
  for (;;) {
    int min_c = -1;
    double min_alpha = 1.0;
    for (int i = 0; i < w->nc; ++i) {
      Contact* c = &w->c[i];
      if (!c->enabled || c->toi_count > MAX_SUB_STEPS) continue;
      double alpha = 1.0;
      if (c->toi_flag) {
        alpha = c->toi;
      } else {
        Fix *FA = &w->f[c->fa], *FB = &w->f[c->fb];
        if (FA->sensor || FB->sensor) continue;
        Body *A = &w->b[FA->body], *B = &w->b[FB->body];
        int active_a = A->awake && A->type != B2G_STATIC, active_b = B->awake && B->type != B2G_STATIC;
        if (!active_a && !active_b) continue;
        int collide_a = A->bullet || A->type != B2G_DYNAMIC, collide_b = B->bullet || B->type != B2G_DYNAMIC;
        if (!collide_a && !collide_b) continue;
        double alpha0 = A->alpha0;
        if (A->alpha0 < B->alpha0) {
          alpha0 = B->alpha0;
          SweepT sw = body_sweep(A); sweep_advance(&sw, alpha0); set_body_sweep(A, &sw);
        } else if (B->alpha0 < A->alpha0) {
          SweepT sw = body_sweep(B); sweep_advance(&sw, alpha0); set_body_sweep(B, &sw);
        }
        DProxy pa, pb;
        dproxy_init(&pa, &FA->sh);
        dproxy_init(&pb, &FB->sh);
        SweepT sa = body_sweep(A), sb = body_sweep(B);
        double beta;
        int state = time_of_impact(w, &pa, &pb, &sa, &sb, &beta);
        alpha = state == TOI_TOUCHING ? pmin(alpha0 + (1.0 - alpha0) * beta, 1.0) : 1.0;
        c->toi = alpha;
        c->toi_flag = 1;
      }
      if (alpha < min_alpha) { min_c = i; min_alpha = alpha; }
    }
    if (min_c < 0 || (1.0 - 10.0 * EPSILON) < min_alpha) break;
    Contact* mc = &w->c[min_c];
    int ba = w->f[mc->fa].body, bb = w->f[mc->fb].body;
    Body *A = &w->b[ba], *B = &w->b[bb];
    SweepT backup_a = body_sweep(A), backup_b = body_sweep(B);
    body_advance(A, min_alpha);
    body_advance(B, min_alpha);
    contact_update(w, mc, min_c);
    mc->toi_flag = 0;
    mc->toi_count += 1;
    if (!mc->enabled || !mc->touching) {
      mc->enabled = 0;
      set_body_sweep(A, &backup_a); A->xf = xf_from(A->c, A->a, A->lc);
      set_body_sweep(B, &backup_b); B->xf = xf_from(B->c, B->a, B->lc);
      continue;
    }
    set_awake(A, 1);
    set_awake(B, 1);
    is.nbodies = 0;
    is.ncontacts = 0;
    A->island_index = is.nbodies; is.bodies[is.nbodies++] = ba;
    B->island_index = is.nbodies; is.bodies[is.nbodies++] = bb;
    is.contacts[is.ncontacts++] = min_c;
    A->toi_flag = 1; B->toi_flag = 1; mc->toi_island = 1;
    for (int side = 0; side < 2; ++side) {
      int bi = side == 0 ? ba : bb;
      Body* body = &w->b[bi];
      if (body->type != B2G_DYNAMIC) continue;
      for (int k = 0; k < w->nc; ++k) { /* body._contacts, in creation order */
        Contact* c = &w->c[k];
        if (w->f[c->fa].body != bi && w->f[c->fb].body != bi) continue;
        if (is.nbodies == 2 * MAX_TOI_CONTACTS) break; /* island body capacity */
        if (is.ncontacts == MAX_TOI_CONTACTS) break;   /* island contact capacity */
        if (c->toi_island) continue;
        int oi = other_body(w, c, bi);
        Body* other = &w->b[oi];
        if (other->type == B2G_DYNAMIC && !body->bullet && !other->bullet) continue;
        if (w->f[c->fa].sensor || w->f[c->fb].sensor) continue;
        SweepT backup = body_sweep(other);
        if (!other->toi_flag) body_advance(other, min_alpha);
        contact_update(w, c, k);
        if (!c->enabled || !c->touching) {
          set_body_sweep(other, &backup);
          other->xf = xf_from(other->c, other->a, other->lc);
          continue;
        }
        c->toi_island = 1;
        is.contacts[is.ncontacts++] = k;
        if (other->toi_flag) continue;
        other->toi_flag = 1;
        if (other->type != B2G_STATIC) set_awake(other, 1);
        other->island_index = is.nbodies;
        is.bodies[is.nbodies++] = oi;
      }
    }
    Step sub;
    sub.dt = (1.0 - min_alpha) * st->dt;
    sub.inv_dt = 1.0 / sub.dt;
    sub.dt_ratio = 1.0;
    sub.pos_iters = 20;
    sub.vel_iters = st->vel_iters;
    sub.warm_starting = 0;
    w->toi_events++;
    island_solve_toi(w, &is, &sub, A->island_index, B->island_index);
    for (int i = 0; i < is.nbodies; ++i) {
      int bi = is.bodies[i];
      Body* b = &w->b[bi];
      b->toi_flag = 0;
      if (b->type != B2G_DYNAMIC) continue;
      if (sync_body_fixtures(w, bi)) goto out;
      for (int k = 0; k < w->nc; ++k) {
        Contact* c = &w->c[k];
        if (w->f[c->fa].body == bi || w->f[c->fb].body == bi) { c->toi_flag = 0; c->toi_island = 0; }
      }
    }
    find_new_contacts(w); /* new contacts start disabled (Contact.__init__): no TOI until updated */
  }
out:
  w->toi_mode = 0;
  free(is.bodies); free(is.contacts); free(is.pv); free(is.vc); free(is.pc);
}

/* World.step (world.py:307-354) in two parts: up to and with collide, and the rest, so that
   a listener's pre_solve verdicts (Contact.enabled = False) can be applied in between. */
static void world_step_begin(World* w, double dt, int vel_iters, int pos_iters) {
  w->ev_n = 0;
  if (w->new_fixture) {
    find_new_contacts(w);
    w->new_fixture = 0;
  }
  collide(w);
  w->split = 1;
  w->split_dt = dt;
  w->split_vi = vel_iters;
  w->split_pi = pos_iters;
}

static void world_step_end(World* w) {
  double dt = w->split_dt;
  w->split = 0;
  Step st;
  st.dt = dt;
  st.inv_dt = dt > 0.0 ? 1.0 / dt : 0.0;
  st.dt_ratio = w->inv_dt0 * dt;
  st.vel_iters = w->split_vi;
  st.pos_iters = w->split_pi;
  st.warm_starting = w->warm_starting;
  if (st.dt > 0.0) solve(w, &st);
  w->toi_calls = 0;
  w->toi_events = 0;
  if (w->continuous && st.dt > 0.0) solve_toi(w, &st);
  if (st.dt > 0.0) w->inv_dt0 = st.inv_dt;
  if (!w->keep_forces)
    for (int i = 0; i < w->nb; ++i) { w->b[i].force = v2(0.0, 0.0); w->b[i].torque = 0.0; }
}

static void world_step(World* w, double dt, int vel_iters, int pos_iters) { /* world.py:307-354 */
  world_step_begin(w, dt, vel_iters, pos_iters);
  world_step_end(w);
}

/* ------------------------------------------------------------------ */
/* load / save                                                         */
/* ------------------------------------------------------------------ */
static void shape_load(Shape* s, const double* d, const int32_t* fi) {
  memset(s, 0, sizeof(*s));
  s->type = fi[B2G_FI_SHAPE];
  s->radius = d[B2G_FD_RADIUS];
  s->n = fi[B2G_FI_NVERTS];
  for (int i = 0; i < B2G_MAX_POLY_VERTS; ++i) {
    s->v[i] = v2(d[B2G_FD_VERTS + 2 * i], d[B2G_FD_VERTS + 2 * i + 1]);
    s->nrm[i] = v2(d[B2G_FD_NORMALS + 2 * i], d[B2G_FD_NORMALS + 2 * i + 1]);
  }
  s->centroid = v2(d[B2G_FD_CENTROID_X], d[B2G_FD_CENTROID_Y]);
  s->p = v2(d[B2G_FD_CIRCLE_PX], d[B2G_FD_CIRCLE_PY]);
  s->e0 = v2(d[B2G_FD_EDGE_V0X], d[B2G_FD_EDGE_V0Y]);
  s->e1 = v2(d[B2G_FD_EDGE_V1X], d[B2G_FD_EDGE_V1Y]);
  s->e2 = v2(d[B2G_FD_EDGE_V2X], d[B2G_FD_EDGE_V2Y]);
  s->e3 = v2(d[B2G_FD_EDGE_V3X], d[B2G_FD_EDGE_V3Y]);
  s->has0 = (fi[B2G_FI_EDGE_FLAGS] & B2G_EF_HAS_V0) != 0;
  s->has3 = (fi[B2G_FI_EDGE_FLAGS] & B2G_EF_HAS_V3) != 0;
}

static void manifold_load(Manifold* m, const double* d, const int32_t* ci) {
  memset(m, 0, sizeof(*m));
  m->type = ci[B2G_CI_MTYPE];
  m->pc = ci[B2G_CI_PC];
  m->ln = v2(d[B2G_CD_LN_X], d[B2G_CD_LN_Y]);
  m->lp = v2(d[B2G_CD_LP_X], d[B2G_CD_LP_Y]);
  m->pts[0].lp = v2(d[B2G_CD_P0_X], d[B2G_CD_P0_Y]);
  m->pts[0].ni = d[B2G_CD_P0_NI];
  m->pts[0].ti = d[B2G_CD_P0_TI];
  m->pts[1].lp = v2(d[B2G_CD_P1_X], d[B2G_CD_P1_Y]);
  m->pts[1].ni = d[B2G_CD_P1_NI];
  m->pts[1].ti = d[B2G_CD_P1_TI];
  mp_id_unpack(ci[B2G_CI_ID0], m->pts[0].id);
  mp_id_unpack(ci[B2G_CI_ID1], m->pts[1].id);
}

static void manifold_save(const Manifold* m, double* d, int32_t* ci) {
  ci[B2G_CI_MTYPE] = m->type;
  ci[B2G_CI_PC] = m->pc;
  d[B2G_CD_LN_X] = m->ln.x; d[B2G_CD_LN_Y] = m->ln.y;
  d[B2G_CD_LP_X] = m->lp.x; d[B2G_CD_LP_Y] = m->lp.y;
  d[B2G_CD_P0_X] = m->pts[0].lp.x; d[B2G_CD_P0_Y] = m->pts[0].lp.y;
  d[B2G_CD_P0_NI] = m->pts[0].ni; d[B2G_CD_P0_TI] = m->pts[0].ti;
  d[B2G_CD_P1_X] = m->pts[1].lp.x; d[B2G_CD_P1_Y] = m->pts[1].lp.y;
  d[B2G_CD_P1_NI] = m->pts[1].ni; d[B2G_CD_P1_TI] = m->pts[1].ti;
  ci[B2G_CI_ID0] = mp_id_pack(m->pts[0].id);
  ci[B2G_CI_ID1] = mp_id_pack(m->pts[1].id);
}

static void world_free(World* w) {
  free(w->b); free(w->f); free(w->c); free(w->move);
  free(w->bfix_off); free(w->bfix); free(w->hint);
  free(w->ev_i); free(w->ev_d); free(w->accept); free(w->slot_map); free(w->new_rows);
  free(w->j); free(w->bj_off); free(w->bj);
  memset(w, 0, sizeof(*w));
}

static int world_load(World* w, const b2g_batch* d, int wi) {
  memset(w, 0, sizeof(*w));
  int b0 = d->body_off[wi], b1 = d->body_off[wi + 1];
  int f0 = d->fix_off[wi], f1 = d->fix_off[wi + 1];
  int c0 = d->con_off[wi], c1 = d->con_off[wi + 1];
  int m0 = d->move_off[wi], m1 = d->move_off[wi + 1];
  const double* wd = d->world_d + (size_t)wi * B2G_WORLD_ND;
  int wf = d->world_i[(size_t)wi * B2G_WORLD_NI + B2G_WI_FLAGS];
  w->gravity = v2(wd[B2G_WD_GX], wd[B2G_WD_GY]);
  w->inv_dt0 = wd[B2G_WD_INV_DT0];
  w->allow_sleep = (wf & B2G_WF_ALLOW_SLEEP) != 0;
  w->warm_starting = (wf & B2G_WF_WARM_STARTING) != 0;
  w->new_fixture = (wf & B2G_WF_NEW_FIXTURE) != 0;
  w->keep_forces = (wf & 8) != 0;
  w->continuous = (wf & B2G_WF_CONTINUOUS) != 0;
  w->nb = b1 - b0;
  w->nf = f1 - f0;
  w->nc = c1 - c0;
  w->nmove = m1 - m0;
  w->b = (Body*)calloc((size_t)w->nb + 1, sizeof(Body));
  w->f = (Fix*)calloc((size_t)w->nf + 1, sizeof(Fix));
  w->ccap = w->nc + 16;
  w->c = (Contact*)calloc((size_t)w->ccap, sizeof(Contact));
  w->mcap = w->nmove + w->nf + 16;
  w->move = (int*)calloc((size_t)w->mcap, sizeof(int));
  w->bfix_off = (int*)calloc((size_t)w->nb + 1, sizeof(int));
  w->bfix = (int*)calloc((size_t)w->nf + 1, sizeof(int));
  if (!w->b || !w->f || !w->c || !w->move || !w->bfix_off || !w->bfix) return -1;
  for (int i = 0; i < w->nb; ++i) {
    const double* bd = d->body_d + (size_t)(b0 + i) * B2G_BODY_ND;
    const int32_t* bi = d->body_i + (size_t)(b0 + i) * B2G_BODY_NI;
    Body* b = &w->b[i];
    b->type = bi[B2G_BI_TYPE];
    int fl = bi[B2G_BI_FLAGS];
    b->awake = (fl & B2G_BF_AWAKE) != 0;
    b->allow_sleep = (fl & B2G_BF_ALLOW_SLEEP) != 0;
    b->fixed_rot = (fl & B2G_BF_FIXED_ROT) != 0;
    b->active = (fl & B2G_BF_ACTIVE) != 0;
    b->bullet = (fl & B2G_BF_BULLET) != 0;
    b->island_id = -1;
    b->xf.p = v2(bd[B2G_BD_XF_PX], bd[B2G_BD_XF_PY]);
    b->xf.R.c1x = bd[B2G_BD_XF_C]; b->xf.R.c1y = bd[B2G_BD_XF_S];
    b->xf.R.c2x = -bd[B2G_BD_XF_S]; b->xf.R.c2y = bd[B2G_BD_XF_C];
    b->lc = v2(bd[B2G_BD_LC_X], bd[B2G_BD_LC_Y]);
    b->c0 = v2(bd[B2G_BD_C0_X], bd[B2G_BD_C0_Y]);
    b->c = v2(bd[B2G_BD_C_X], bd[B2G_BD_C_Y]);
    b->a0 = bd[B2G_BD_A0]; b->a = bd[B2G_BD_A];
    b->v = v2(bd[B2G_BD_V_X], bd[B2G_BD_V_Y]);
    b->w = bd[B2G_BD_W];
    b->force = v2(bd[B2G_BD_F_X], bd[B2G_BD_F_Y]);
    b->torque = bd[B2G_BD_TORQUE];
    b->mass = bd[B2G_BD_MASS]; b->inv_mass = bd[B2G_BD_INV_MASS];
    b->I = bd[B2G_BD_I]; b->invI = bd[B2G_BD_INV_I];
    b->gscale = bd[B2G_BD_GRAV_SCALE];
    b->ldamp = bd[B2G_BD_LIN_DAMP]; b->adamp = bd[B2G_BD_ANG_DAMP];
    b->sleep_time = bd[B2G_BD_SLEEP_TIME];
  }
  for (int i = 0; i < w->nf; ++i) {
    const double* fd = d->fix_d + (size_t)(f0 + i) * B2G_FIX_ND;
    const int32_t* fi = d->fix_i + (size_t)(f0 + i) * B2G_FIX_NI;
    Fix* f = &w->f[i];
    f->body = fi[B2G_FI_BODY];
    if (f->body < 0 || f->body >= w->nb) return -2;
    shape_load(&f->sh, fd, fi);
    if (f->sh.type == B2G_SHAPE_LOOP || f->sh.n > B2G_MAX_POLY_VERTS) return -3;
    f->friction = fd[B2G_FD_FRICTION];
    f->restitution = fd[B2G_FD_RESTITUTION];
    f->density = fd[B2G_FD_DENSITY];
    f->category = fi[B2G_FI_CATEGORY];
    f->mask = fi[B2G_FI_MASK];
    f->group = fi[B2G_FI_GROUP];
    f->sensor = fi[B2G_FI_SENSOR];
    f->fat.lo = v2(fd[B2G_FD_FAT_LX], fd[B2G_FD_FAT_LY]);
    f->fat.hi = v2(fd[B2G_FD_FAT_UX], fd[B2G_FD_FAT_UY]);
    w->bfix_off[f->body + 1]++;
  }
  for (int i = 0; i < w->nb; ++i) w->bfix_off[i + 1] += w->bfix_off[i];
  {
    int* fill = (int*)calloc((size_t)w->nb + 1, sizeof(int));
    if (!fill) return -1;
    for (int i = 0; i < w->nf; ++i) {
      int b = w->f[i].body;
      w->bfix[w->bfix_off[b] + fill[b]++] = i;
    }
    free(fill);
  }
  for (int i = 0; i < w->nc; ++i) {
    const double* cd = d->con_d + (size_t)(c0 + i) * B2G_CON_ND;
    const int32_t* ci = d->con_i + (size_t)(c0 + i) * B2G_CON_NI;
    Contact* c = &w->c[i];
    c->fa = ci[B2G_CI_FA];
    c->fb = ci[B2G_CI_FB];
    if (c->fa < 0 || c->fa >= w->nf || c->fb < 0 || c->fb >= w->nf) return -2;
    int fl = ci[B2G_CI_FLAGS];
    c->touching = (fl & B2G_CF_TOUCHING) != 0;
    c->enabled = (fl & B2G_CF_ENABLED) != 0;
    c->filter = (fl & B2G_CF_FILTER) != 0;
    c->friction = cd[B2G_CD_FRICTION];
    c->restitution = cd[B2G_CD_RESTITUTION];
    manifold_load(&c->m, cd, ci);
  }
  for (int i = 0; i < w->nmove; ++i) w->move[i] = d->move[m0 + i];
  /* joints, and each body's joint list (world.py:268-273) */
  int j0 = d->joint_off ? d->joint_off[wi] : 0, j1 = d->joint_off ? d->joint_off[wi + 1] : 0;
  w->nj = j1 - j0;
  w->j = (Joint*)calloc((size_t)w->nj + 1, sizeof(Joint));
  w->bj_off = (int*)calloc((size_t)w->nb + 1, sizeof(int));
  w->bj = (int*)calloc((size_t)(2 * w->nj) + 1, sizeof(int));
  if (!w->j || !w->bj_off || !w->bj) return -1;
  for (int k = 0; k < w->nj; ++k) {
    Joint* J = &w->j[k];
    memcpy(J->i, d->joint_i + (size_t)(j0 + k) * B2G_JOINT_NI, sizeof(int32_t) * B2G_JOINT_NI);
    memcpy(J->d, d->joint_d + (size_t)(j0 + k) * B2G_JOINT_ND, sizeof(double) * B2G_JOINT_ND);
    int t = J->i[B2G_JI_TYPE], a = J->i[B2G_JI_BODY_A], b = J->i[B2G_JI_BODY_B];
    if (t < B2G_JT_REVOLUTE || t > B2G_JT_ROPE) return -4;
    if (a < -1 || a >= w->nb || b < 0 || b >= w->nb || (a < 0 && t != B2G_JT_MOUSE)) return -2;
    if (t == B2G_JT_GEAR) {
      int c = J->i[B2G_JI_BODY_C], dd = J->i[B2G_JI_BODY_D];
      if (a < 0 || c < 0 || c >= w->nb || dd < 0 || dd >= w->nb) return -2;
    }
    if (a >= 0) w->bj_off[a + 1]++;
    w->bj_off[b + 1]++;
  }
  for (int i = 0; i < w->nb; ++i) w->bj_off[i + 1] += w->bj_off[i];
  {
    int* fill = (int*)calloc((size_t)w->nb + 1, sizeof(int));
    if (!fill) return -1;
    for (int k = 0; k < w->nj; ++k) {
      int a = w->j[k].i[B2G_JI_BODY_A], b = w->j[k].i[B2G_JI_BODY_B];
      if (a >= 0) w->bj[w->bj_off[a] + fill[a]++] = k;
      w->bj[w->bj_off[b] + fill[b]++] = k;
    }
    free(fill);
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* exported entry points (ctypes)                                      */
/* ------------------------------------------------------------------ */

ORC_EXPORT void* orc_load(const b2g_batch* d) {
  init_consts();
  Batch* B = (Batch*)calloc(1, sizeof(Batch));
  if (!B) return NULL;
  B->nw = d->n_worlds;
  B->w = (World*)calloc((size_t)(B->nw > 0 ? B->nw : 1), sizeof(World));
  if (!B->w) { free(B); return NULL; }
  for (int i = 0; i < B->nw; ++i) {
    if (world_load(&B->w[i], d, i)) {
      for (int j = 0; j <= i; ++j) world_free(&B->w[j]);
      free(B->w);
      free(B);
      return NULL;
    }
    B->w[i].index = i;
  }
  return B;
}

ORC_EXPORT void orc_free(void* h) {
  Batch* B = (Batch*)h;
  if (!B) return;
  for (int i = 0; i < B->nw; ++i) world_free(&B->w[i]);
  free(B->w);
  free(B);
}

/* Steps every world once; returns the OR of the worlds' error words. */
ORC_EXPORT int orc_step(void* h, double dt, int vel_iters, int pos_iters) {
  Batch* B = (Batch*)h;
  int err = 0;
  for (int i = 0; i < B->nw; ++i) {
    world_step(&B->w[i], dt, vel_iters, pos_iters);
    err |= B->w[i].err;
  }
  return err;
}

ORC_EXPORT int orc_step_world(void* h, int wi, double dt, int vel_iters, int pos_iters) {
  Batch* B = (Batch*)h;
  if (wi < 0 || wi >= B->nw) return -1;
  world_step(&B->w[wi], dt, vel_iters, pos_iters);
  return B->w[wi].err;
}

/* counts[0..3] = total bodies, fixtures, contacts, move-buffer entries */
ORC_EXPORT void orc_counts(void* h, int32_t* counts) {
  Batch* B = (Batch*)h;
  int32_t nb = 0, nf = 0, nc = 0, nm = 0;
  for (int i = 0; i < B->nw; ++i) { nb += B->w[i].nb; nf += B->w[i].nf; nc += B->w[i].nc; nm += B->w[i].nmove; }
  counts[0] = nb; counts[1] = nf; counts[2] = nc; counts[3] = nm;
}

ORC_EXPORT int orc_save(void* h, b2g_batch_out* o) {
  Batch* B = (Batch*)h;
  if (o->n_worlds != B->nw) return -1;
  int bo = 0, fo = 0, co = 0, mo = 0, jo = 0;
  o->con_off[0] = 0;
  o->move_off[0] = 0;
  for (int wi = 0; wi < B->nw; ++wi) {
    World* w = &B->w[wi];
    double* wd = o->world_d + (size_t)wi * B2G_WORLD_ND;
    int32_t* wfl = o->world_i + (size_t)wi * B2G_WORLD_NI;
    wd[B2G_WD_GX] = w->gravity.x; wd[B2G_WD_GY] = w->gravity.y; wd[B2G_WD_INV_DT0] = w->inv_dt0;
    wfl[B2G_WI_FLAGS] = (w->allow_sleep ? B2G_WF_ALLOW_SLEEP : 0) | (w->warm_starting ? B2G_WF_WARM_STARTING : 0) |
                        (w->new_fixture ? B2G_WF_NEW_FIXTURE : 0) | (w->keep_forces ? 8 : 0) |
                        (w->continuous ? B2G_WF_CONTINUOUS : 0);
    for (int i = 0; i < w->nb; ++i) {
      double* bd = o->body_d + (size_t)(bo + i) * B2G_BODY_ND;
      int32_t* bi = o->body_i + (size_t)(bo + i) * B2G_BODY_NI;
      const Body* b = &w->b[i];
      memset(bd, 0, sizeof(double) * B2G_BODY_ND);
      bi[B2G_BI_TYPE] = b->type;
      bi[B2G_BI_FLAGS] = (b->awake ? B2G_BF_AWAKE : 0) | (b->allow_sleep ? B2G_BF_ALLOW_SLEEP : 0) |
                         (b->fixed_rot ? B2G_BF_FIXED_ROT : 0) | (b->active ? B2G_BF_ACTIVE : 0) |
                         (b->bullet ? B2G_BF_BULLET : 0);
      bi[B2G_BI_ISLAND] = b->island_id;
      bi[B2G_BI_PAD] = 0;
      bd[B2G_BD_XF_PX] = b->xf.p.x; bd[B2G_BD_XF_PY] = b->xf.p.y;
      bd[B2G_BD_XF_C] = b->xf.R.c1x; bd[B2G_BD_XF_S] = b->xf.R.c1y;
      bd[B2G_BD_LC_X] = b->lc.x; bd[B2G_BD_LC_Y] = b->lc.y;
      bd[B2G_BD_C0_X] = b->c0.x; bd[B2G_BD_C0_Y] = b->c0.y;
      bd[B2G_BD_C_X] = b->c.x; bd[B2G_BD_C_Y] = b->c.y;
      bd[B2G_BD_A0] = b->a0; bd[B2G_BD_A] = b->a;
      bd[B2G_BD_V_X] = b->v.x; bd[B2G_BD_V_Y] = b->v.y; bd[B2G_BD_W] = b->w;
      bd[B2G_BD_F_X] = b->force.x; bd[B2G_BD_F_Y] = b->force.y; bd[B2G_BD_TORQUE] = b->torque;
      bd[B2G_BD_MASS] = b->mass; bd[B2G_BD_INV_MASS] = b->inv_mass;
      bd[B2G_BD_I] = b->I; bd[B2G_BD_INV_I] = b->invI;
      bd[B2G_BD_GRAV_SCALE] = b->gscale; bd[B2G_BD_LIN_DAMP] = b->ldamp; bd[B2G_BD_ANG_DAMP] = b->adamp;
      bd[B2G_BD_SLEEP_TIME] = b->sleep_time;
    }
    for (int i = 0; i < w->nf; ++i) {
      double* fd = o->fix_d + (size_t)(fo + i) * B2G_FIX_ND;
      const Fix* f = &w->f[i];
      fd[B2G_FD_FAT_LX] = f->fat.lo.x; fd[B2G_FD_FAT_LY] = f->fat.lo.y;
      fd[B2G_FD_FAT_UX] = f->fat.hi.x; fd[B2G_FD_FAT_UY] = f->fat.hi.y;
    }
    if (co + w->nc > o->con_cap || mo + w->nmove > o->move_cap) return -2;
    for (int i = 0; i < w->nc; ++i) {
      double* cd = o->con_d + (size_t)(co + i) * B2G_CON_ND;
      int32_t* ci = o->con_i + (size_t)(co + i) * B2G_CON_NI;
      const Contact* c = &w->c[i];
      memset(cd, 0, sizeof(double) * B2G_CON_ND);
      memset(ci, 0, sizeof(int32_t) * B2G_CON_NI);
      ci[B2G_CI_FA] = c->fa; ci[B2G_CI_FB] = c->fb;
      ci[B2G_CI_FLAGS] = (c->touching ? B2G_CF_TOUCHING : 0) | (c->enabled ? B2G_CF_ENABLED : 0) |
                         (c->filter ? B2G_CF_FILTER : 0);
      cd[B2G_CD_FRICTION] = c->friction; cd[B2G_CD_RESTITUTION] = c->restitution;
      manifold_save(&c->m, cd, ci);
    }
    for (int i = 0; i < w->nmove; ++i) o->move[mo + i] = w->move[i];
    if (o->joint_d && o->joint_i)
      for (int k = 0; k < w->nj; ++k) {
        memcpy(o->joint_i + (size_t)(jo + k) * B2G_JOINT_NI, w->j[k].i, sizeof(int32_t) * B2G_JOINT_NI);
        memcpy(o->joint_d + (size_t)(jo + k) * B2G_JOINT_ND, w->j[k].d, sizeof(double) * B2G_JOINT_ND);
      }
    jo += w->nj;
    bo += w->nb; fo += w->nf; co += w->nc; mo += w->nmove;
    o->con_off[wi + 1] = co;
    o->move_off[wi + 1] = mo;
  }
  return 0;
}

/* The solver seam: the island search and Island.solve of every world on its current state
   (no collide, no fixture synchronisation or pair search, forces kept); returns the OR of the
   worlds' error words. */
ORC_EXPORT int orc_solve_islands(void* h, double dt, int vel_iters, int pos_iters) {
  Batch* B = (Batch*)h;
  int err = 0;
  for (int i = 0; i < B->nw; ++i) {
    World* w = &B->w[i];
    Step st;
    st.dt = dt;
    st.inv_dt = dt > 0.0 ? 1.0 / dt : 0.0;
    st.dt_ratio = w->inv_dt0 * dt;
    st.vel_iters = vel_iters;
    st.pos_iters = pos_iters;
    st.warm_starting = w->warm_starting;
    w->ev_n = 0;
    if (dt > 0.0) solve_islands(w, &st, 0);
    err |= w->err;
  }
  return err;
}

/* Replay hint for world wi: n unordered fixture pairs (2n ints) in the
   reference's contact-creation order.  n = 0 removes the hint. */
ORC_EXPORT int orc_set_hint(void* h, int wi, const int32_t* pairs, int n) {
  Batch* B = (Batch*)h;
  if (wi < 0 || wi >= B->nw) return -1;
  World* w = &B->w[wi];
  free(w->hint);
  w->hint = NULL;
  w->hint_n = 0;
  if (n <= 0) return 0;
  w->hint = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
  if (!w->hint) return -1;
  for (int i = 0; i < n; ++i) w->hint[i] = ukey(pairs[2 * i], pairs[2 * i + 1]);
  w->hint_n = n;
  return 0;
}

/* World wi's step in two parts: orc_step_begin (the pending pair search and collide), then
   orc_disable for the contacts a pre_solve disabled (slot: the contact's index during collide,
   a phase-0 record's key / 4), then orc_step_end.  Each returns the world's error word, or -1. */
ORC_EXPORT int orc_step_begin(void* h, int wi, double dt, int vel_iters, int pos_iters) {
  Batch* B = (Batch*)h;
  if (wi < 0 || wi >= B->nw || B->w[wi].split) return -1;
  world_step_begin(&B->w[wi], dt, vel_iters, pos_iters);
  return B->w[wi].err;
}

ORC_EXPORT int orc_disable(void* h, int wi, int slot) {
  Batch* B = (Batch*)h;
  if (wi < 0 || wi >= B->nw) return -1;
  World* w = &B->w[wi];
  if (!w->split || slot < 0 || slot >= w->slot_n || w->slot_map[slot] < 0) return -1;
  w->c[w->slot_map[slot]].enabled = 0;
  return 0;
}

ORC_EXPORT int orc_step_end(void* h, int wi) {
  Batch* B = (Batch*)h;
  if (wi < 0 || wi >= B->nw || !B->w[wi].split) return -1;
  world_step_end(&B->w[wi]);
  return B->w[wi].err;
}

/* World wi's continuous-physics counts of its last step: time_of_impact calls and TOI events
   (sub-step island solves). */
ORC_EXPORT int orc_toi_stats(void* h, int wi, int64_t* out) {
  Batch* B = (Batch*)h;
  if (wi < 0 || wi >= B->nw) return -1;
  out[0] = B->w[wi].toi_calls;
  out[1] = B->w[wi].toi_events;
  return 0;
}

/* The contact listener kinds to record (1 << B2G_EV_*), for every world. */
ORC_EXPORT void orc_set_events(void* h, int mask) {
  Batch* B = (Batch*)h;
  for (int i = 0; i < B->nw; ++i) B->w[i].ev_mask = mask;
}

/* The last step's listener records of world wi, in call order.  Returns the count,
   -1 (bad world) or -2 (more than cap records: nothing written). */
ORC_EXPORT int orc_events(void* h, int wi, int32_t* out_i, double* out_d, int cap) {
  Batch* B = (Batch*)h;
  if (wi < 0 || wi >= B->nw) return -1;
  const World* w = &B->w[wi];
  if (w->ev_n > cap) return -2;
  if (w->ev_n > 0) {
    memcpy(out_i, w->ev_i, sizeof(int32_t) * B2G_EV_NI * (size_t)w->ev_n);
    memcpy(out_d, w->ev_d, sizeof(double) * B2G_EV_ND * (size_t)w->ev_n);
  }
  return w->ev_n;
}

/* A custom contact_filter for world wi as its verdicts: accept[a * nf + b] != 0 when
   filter(fixture a, fixture b) returns True; accept == NULL restores the default filter. */
ORC_EXPORT int orc_set_filter(void* h, int wi, const unsigned char* accept, int nf) {
  Batch* B = (Batch*)h;
  if (wi < 0 || wi >= B->nw) return -1;
  World* w = &B->w[wi];
  free(w->accept);
  w->accept = NULL;
  if (!accept) return 0;
  if (nf != w->nf) return -1;
  w->accept = (unsigned char*)malloc((size_t)nf * (size_t)nf + 1);
  if (!w->accept) return -1;
  memcpy(w->accept, accept, (size_t)nf * (size_t)nf);
  return 0;
}

/* The contacts world wi's last find_new_contacts created, 3 ints each: fixture A, fixture B,
   B2G_CF_SWAPPED when the registry swapped the reported pair.  Returns the count or -2. */
ORC_EXPORT int orc_new_contacts(void* h, int wi, int32_t* out, int cap) {
  Batch* B = (Batch*)h;
  if (wi < 0 || wi >= B->nw) return -1;
  const World* w = &B->w[wi];
  if (w->new_n > cap) return -2;
  if (w->new_n > 0) memcpy(out, w->new_rows, sizeof(int32_t) * 3 * (size_t)w->new_n);
  return w->new_n;
}

ORC_EXPORT int orc_error(void* h) {
  Batch* B = (Batch*)h;
  int e = 0;
  for (int i = 0; i < B->nw; ++i) e |= B->w[i].err;
  return e;
}

/* Narrow-phase seam: evaluates one contact (fixtures A, B oriented as the
   registry orders them) at body transforms xf = (px, py, c, s), starting from
   a zeroed manifold.  Writes the manifold into one contact record. */
ORC_EXPORT int orc_evaluate(const double* fa_d, const int32_t* fa_i, const double* fb_d, const int32_t* fb_i,
                            const double* xfa, const double* xfb, double* con_d, int32_t* con_i) {
  init_consts();
  Shape A, B;
  shape_load(&A, fa_d, fa_i);
  shape_load(&B, fb_d, fb_i);
  XF ta, tb;
  ta.p = v2(xfa[0], xfa[1]); ta.R.c1x = xfa[2]; ta.R.c1y = xfa[3]; ta.R.c2x = -xfa[3]; ta.R.c2y = xfa[2];
  tb.p = v2(xfb[0], xfb[1]); tb.R.c1x = xfb[2]; tb.R.c1y = xfb[3]; tb.R.c2x = -xfb[3]; tb.R.c2y = xfb[2];
  Manifold m;
  memset(&m, 0, sizeof(m));
  if (evaluate(&m, &A, ta, &B, tb)) return -1;
  manifold_save(&m, con_d, con_i);
  return 0;
}

/* Broad-phase seam: pairs among n fat AABBs (n x 4: lx, ly, ux, uy) for the
   move buffer `move`, in the oracle's canonical order (move-buffer order, then
   ascending partner).  Returns the pair count, or -1 (allocation) / -2 (cap). */
ORC_EXPORT int orc_pairs(int n, const double* fat, int nmove, const int32_t* move, int32_t* out, int cap) {
  init_consts();
  AABB* a = (AABB*)malloc(sizeof(AABB) * (size_t)(n > 0 ? n : 1));
  int* mv = (int*)malloc(sizeof(int) * (size_t)(nmove > 0 ? nmove : 1));
  if (!a || !mv) { free(a); free(mv); return -1; }
  for (int i = 0; i < n; ++i) { a[i].lo = v2(fat[4 * i], fat[4 * i + 1]); a[i].hi = v2(fat[4 * i + 2], fat[4 * i + 3]); }
  for (int i = 0; i < nmove; ++i) mv[i] = move[i];
  int* pairs = NULL;
  int np = 0;
  int rc = update_pairs(a, n, mv, nmove, &pairs, &np);
  free(a);
  free(mv);
  if (rc) return -1;
  if (np > cap) { free(pairs); return -2; }
  for (int i = 0; i < 2 * np; ++i) out[i] = pairs[i];
  free(pairs);
  return np;
}

/* AABB of one fixture's shape at a transform (px, py, c, s); out = lx, ly, ux, uy. */
ORC_EXPORT void orc_compute_aabb(const double* fd, const int32_t* fi, const double* xf, double* out) {
  init_consts();
  Shape s;
  shape_load(&s, fd, fi);
  XF t;
  t.p = v2(xf[0], xf[1]); t.R.c1x = xf[2]; t.R.c1y = xf[3]; t.R.c2x = -xf[3]; t.R.c2y = xf[2];
  AABB a = compute_aabb(&s, t);
  out[0] = a.lo.x; out[1] = a.lo.y; out[2] = a.hi.x; out[3] = a.hi.y;
}

/* ------------------------------------------------------------------ */
/* Queries and ray casts: World.query_aabb / World.ray_cast            */
/* (world.py:371-440), by brute force over a world's proxies.          */
/* ------------------------------------------------------------------ */
/* Circle.ray_cast (shapes.py:147-174) */
static int ray_circle(const Shape* sh, XF xf, V2 p1, V2 p2, double maxf, double* frac, V2* normal) {
  V2 position = add(xf.p, mulmv(xf.R, sh->p));
  V2 s = sub(p1, position);
  double b = dot(s, s) - pysq(sh->radius);
  V2 r = sub(p2, p1);
  double c = dot(s, r);
  double rr = dot(r, r);
  double sigma = c * c - rr * b;
  if (sigma < 0.0 || rr < EPSILON) return 0;
  double a = -(c + sqrt(sigma));
  if (0.0 <= a && a <= maxf * rr) {
    a /= rr;
    *frac = a;
    *normal = normalized(add(s, scale(a, r)));
    return 1;
  }
  return 0;
}

/* Polygon.ray_cast (shapes.py:265-307) */
static int ray_polygon(const Shape* sh, XF xf, V2 p1w, V2 p2w, double maxf, double* frac, V2* normal) {
  V2 p1 = mulmtv(xf.R, sub(p1w, xf.p));
  V2 p2 = mulmtv(xf.R, sub(p2w, xf.p));
  V2 d = sub(p2, p1);
  double lower = 0.0, upper = maxf;
  int tagged = -1;
  for (int i = 0; i < sh->n; ++i) {
    double numerator = dot(sh->nrm[i], sub(sh->v[i], p1));
    double denominator = dot(sh->nrm[i], d);
    if (denominator == 0.0) {
      if (numerator < 0.0) return 0;
    } else {
      if (denominator < 0.0 && numerator < lower * denominator) {
        lower = numerator / denominator;
        tagged = i;
      } else if (denominator > 0.0 && numerator < upper * denominator) {
        upper = numerator / denominator;
      }
    }
    if (upper < lower) return 0;
  }
  if (tagged < 0) return 0;
  *normal = mulmv(xf.R, sh->nrm[tagged]);
  *frac = lower;
  return 1;
}

/* Edge.ray_cast (shapes.py:655-702): the normal stays in the edge's frame */
static int ray_edge(const Shape* sh, XF xf, V2 p1w, V2 p2w, double* frac, V2* normal) {
  V2 p1 = mulmtv(xf.R, sub(p1w, xf.p));
  V2 p2 = mulmtv(xf.R, sub(p2w, xf.p));
  V2 d = sub(p2, p1);
  V2 v1 = sh->e1, v2_ = sh->e2;
  V2 e = sub(v2_, v1);
  V2 n = normalized(v2(e.y, -e.x));
  double numerator = dot(n, sub(v1, p1));
  double denominator = dot(n, d);
  if (denominator == 0.0) return 0;
  double t = numerator / denominator;
  if (t < 0.0 || 1.0 < t) return 0;
  V2 q = add(p1, scale(t, d));
  V2 r = sub(v2_, v1);
  double rr = dot(r, r);
  if (rr == 0.0) return 0;
  double sv = dot(sub(q, v1), r) / rr;
  if (sv < 0.0 || 1.0 < sv) return 0;
  *frac = t;
  if (numerator > 0.0) n = neg(n);
  *normal = n;
  return 1;
}

static int shape_ray_cast(const Shape* sh, XF xf, V2 p1, V2 p2, double maxf, double* frac, V2* normal) {
  if (sh->type == B2G_SHAPE_CIRCLE) return ray_circle(sh, xf, p1, p2, maxf, frac, normal);
  if (sh->type == B2G_SHAPE_POLYGON) return ray_polygon(sh, xf, p1, p2, maxf, frac, normal);
  if (sh->type == B2G_SHAPE_EDGE) return ray_edge(sh, xf, p1, p2, frac, normal);
  return 0;
}

/* One shape's ray cast at a transform (px, py, c, s) for seg = (p1 x, p1 y, p2 x, p2 y) and
   max_fraction; out = fraction, normal x, normal y.  Returns 1 on a hit. */
ORC_EXPORT int orc_shape_ray_cast(const double* fd, const int32_t* fi, const double* xf, const double* seg,
                                  double max_fraction, double* out) {
  init_consts();
  Shape sh;
  shape_load(&sh, fd, fi);
  XF t;
  t.p = v2(xf[0], xf[1]); t.R.c1x = xf[2]; t.R.c1y = xf[3]; t.R.c2x = -xf[3]; t.R.c2y = xf[2];
  double f = 0.0;
  V2 n = v2(0.0, 0.0);
  int hit = shape_ray_cast(&sh, t, v2(seg[0], seg[1]), v2(seg[2], seg[3]), max_fraction, &f, &n);
  out[0] = f; out[1] = n.x; out[2] = n.y;
  return hit;
}

/* World.query_aabb: the fixtures of world wi whose fat AABB overlaps box (lx, ly, ux, uy), in
   fixture order (node.aabb.test_overlap(aabb), dynamictree.py:190-207). */
ORC_EXPORT int orc_query_aabb(void* h, int wi, const double* box, int32_t* out, int cap) {
  Batch* B = (Batch*)h;
  if (!B || wi < 0 || wi >= B->nw) return -1;
  World* w = &B->w[wi];
  AABB q;
  q.lo = v2(box[0], box[1]);
  q.hi = v2(box[2], box[3]);
  int n = 0;
  for (int f = 0; f < w->nf; ++f) {
    if (!aabb_overlap(&w->f[f].fat, &q)) continue;
    if (n < cap) out[n] = f;
    ++n;
  }
  return n;
}

/* World.ray_cast with a callback that returns 1 (every hit, no clipping): the fixtures of world
   wi whose proxy passes DynamicTree.ray_cast's test at max_fraction 1 (dynamictree.py:209-273)
   and whose shape the ray hits, in fixture order; out_d = fraction, point x, y, normal x, y
   per hit.  Returns the number of hits, or -1 for p1 == p2 (the reference's ValueError). */
ORC_EXPORT int orc_ray_cast(void* h, int wi, const double* seg, int32_t* out_f, double* out_d, int cap) {
  Batch* B = (Batch*)h;
  if (!B || wi < 0 || wi >= B->nw) return -1;
  init_consts();
  World* w = &B->w[wi];
  V2 p1 = v2(seg[0], seg[1]), p2 = v2(seg[2], seg[3]);
  V2 r = sub(p2, p1);
  if (length_sq(r) <= 0.0) return -1;
  r = normalized(r);
  V2 v = cross_sv(1.0, r);
  V2 abs_v = v2(fabs(v.x), fabs(v.y));
  V2 t = add(p1, scale(1.0, sub(p2, p1)));
  AABB segbox;
  segbox.lo = vmin(p1, t);
  segbox.hi = vmax(p1, t);
  int n = 0;
  for (int f = 0; f < w->nf; ++f) {
    const AABB* fat = &w->f[f].fat;
    if (!aabb_overlap(fat, &segbox)) continue;
    V2 c = scale(0.5, add(fat->lo, fat->hi));
    V2 hh = scale(0.5, sub(fat->hi, fat->lo));
    double separation = fabs(dot(v, sub(p1, c))) - dot(abs_v, hh);
    if (separation > 0.0) continue;
    double fr = 0.0;
    V2 nrm = v2(0.0, 0.0);
    if (!shape_ray_cast(&w->f[f].sh, w->b[w->f[f].body].xf, p1, p2, 1.0, &fr, &nrm)) continue;
    if (n < cap) {
      V2 pt = add(scale(1.0 - fr, p1), scale(fr, p2));
      out_f[n] = f;
      double* o = out_d + 5 * (size_t)n;
      o[0] = fr; o[1] = pt.x; o[2] = pt.y; o[3] = nrm.x; o[4] = nrm.y;
    }
    ++n;
  }
  return n;
}
