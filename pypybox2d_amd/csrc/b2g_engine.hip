// b2g_engine.hip -- batched MI355X (gfx950) contact pipeline and its C-ABI.
//
// One call of b2g_step advances every world of a batch by one
// World.step(dt, velocity_iterations, position_iterations) (world.py:307-354)
// with continuous physics off.  Per step the device runs:
//
//   k_collide   one lane per contact slot: ContactManager.collide
//               (contact_manager.py:174-223) for contacts with an awake
//               non-static body at step start -- fat-AABB overlap test in FP64,
//               manifold evaluation, warm-start id matching (contact.py:210-268).
//   k_islands   one workgroup per world: the rest of collide's sequential
//               awake gating (only worlds holding sleeping bodies), then the
//               island search of World._solve (world.py:494-580) over an
//               LDS-resident contact graph, in the reference's DFS order.
//   k_solve     one lane per island, pulled from a device queue: Island.solve
//               (island.py:65-179) with ContactSolver (contact.py:375-864).
//   k_sync      one lane per fixture: Body._synchronize_fixtures + move_proxy
//               (body.py:898-904, fixture.py:287-298, dynamictree.py:152-188)
//               with FP64 swept AABBs.
//   k_pairs     one workgroup per world: BroadPhase.update_pairs + ContactManager
//               .find_new_contacts (broadphase.py:83-109, contact_manager.py:
//               102-158): exact FP64 fat-AABB tests of every moved proxy against
//               the world's proxies, filtering, and contact creation in a
//               canonical order; then the step's bookkeeping.
//
// The bulk arithmetic (manifolds, effective masses, impulses, velocities) runs
// in T = float (FP32-bulk) or T = double (FP64); positions, angles, transforms,
// fat AABBs and sleep timers are FP64 in both.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/b2g.h"
#include "b2g_collide.hpp"
#include "b2g_raycast.hpp"
#include "b2g_toi.hpp"
#include "b2g_joints.hpp"

namespace b2g {

// contact slot state bits
enum : int {
  CS_LIVE = 1,
  CS_TOUCHING = 2,
  CS_ENABLED = 4,
  CS_FILTER = 8,
  CS_MTYPE_SHIFT = 4,
  CS_PC_SHIFT = 6,
  CS_PROCESSED = 256,
  CS_WOKE = 512,
  CS_SWAPPED = 1024,  // the registry swapped the broad phase's pair at creation
  CS_REJECT = 2048,   // with a custom filter: the host's filter rejected this flagged contact
};
// body flag bits (B2G_BF_*)
enum : int { BF_AWAKE = 1, BF_ALLOW_SLEEP = 2, BF_FIXED_ROT = 4, BF_ACTIVE = 8, BF_BULLET = 16 };
enum : int { BT_STATIC = 0, BT_DYNAMIC = 1, BT_KINEMATIC = 2 };
enum : int {
  WF_ALLOW_SLEEP = 1, WF_WARM = 2, WF_NEW_FIXTURE = 4, WF_KEEP_FORCES = 8, WF_CUSTOM_FILTER = 16, WF_CONTINUOUS = 32
};
// the parts of a step one launch sequence runs: all of it; the new-fixture pair search alone
// (b2g_pair_search); up to and with collide (b2g_step_collide); the rest (b2g_step_solve)
enum : int { PART_ALL = 0, PART_PAIRS = 1, PART_COLLIDE = 2, PART_SOLVE = 3, PART_ISLANDS = 4 };
enum : int {
  WERR_CAPACITY = 1,
  WERR_UNSUPPORTED = 2,
  WERR_NEWPAIRS = 4,
  WERR_SCHEDULE = 8,
  WERR_NAN = 16,          // a body's position, angle or velocity became non-finite
  WERR_ISLAND_SIZE = 32,  // an island of more than 65535 contacts (16-bit island-local indices)
  WERR_TOI = 64,          // a time-of-impact search that the reference would not end (toi.py:335-339)
  WERR_JOINT = 128,       // a mouse joint's spring too weak for its time step (the reference asserts)
};

// work counters (b2g_stats): summed over the steps since the last reset
enum : int {
  S_LIVE = 0, S_EVAL = 1, S_NEW = 2, S_MOVED = 3, S_IBODIES = 4, S_ICONTACTS = 5, S_TOI_CALLS = 6, S_TOI_EVENTS = 7,
  kStats = 8
};

constexpr int kBlock = 256;
constexpr int kBigPairsBlock = 1024;  // k_pairs workgroup for the largest worlds
constexpr int kMaxLds = 163840;  // LDS per workgroup on gfx950
constexpr int kProfileBatch = 128;  // steps per event read-back in b2g_profile_step
constexpr int kSizeClasses = 16;
// q_cls layout: [0, 16) island count per class, [16, 32) cursors,
// [32] islands in team classes, [33] wavefronts for them,
// [34, 50) first wavefront per class, [50, 66) first queue entry per class
constexpr int kQHeavy = 2 * kSizeClasses, kQHeavyWaves = kQHeavy + 1, kQWaveBase = kQHeavy + 2,
              kQBase = kQWaveBase + kSizeClasses, kQcls = kQBase + kSizeClasses;

__host__ __device__ inline int size_class(int nc) {  // floor(log2(nc + 1)), capped
  int c = 0;
  while ((2 << c) <= nc + 1 && c < kSizeClasses - 1) ++c;
  return c;
}

constexpr int kSolveBlock = 64;  // k_solve: one wavefront per workgroup
constexpr int kIslandTimeFields = 6;  // int64 per island of b2g_island_times

// Team size of a heavy size class: 64 lanes from full_cls up, halved per class
// below down to 8, and at most tmax lanes, except in the top class, which holds
// the islands whose state does not fit a smaller team's LDS share (queue_class).
__host__ __device__ inline int team_size(int cls, int full_cls, int tmax) {
  int sh = full_cls - cls;
  sh = sh < 0 ? 0 : (sh > 3 ? 3 : sh);
  const int ts = kSolveBlock >> sh;
  return cls == kSizeClasses - 1 ? kSolveBlock : (ts < tmax ? ts : tmax);
}

// The length of a team island in lane passes per sweep, roughly: a lane walks its sweeps in
// steps of the island's sweep period and owns a team size's share of each sweep's contacts.
__host__ __device__ inline int team_island_length(const int4& in, int nc, int TS) {
  const int period = (int)((uint32_t)in.z >> 16), per_lane = ((nc & 0x3FFFFFFF) + TS - 1) / TS;
  return period > per_lane ? period : per_lane;
}

__host__ __device__ inline int team_lds_bytes(int nbi) {
  return 24 * nbi + 4 * ((nbi + 1) & ~1);  // 3 doubles and an int counter per body
}

// The queue class of an island of nc contacts and nbi bodies: its size class,
// except that a heavy island whose body state does not fit the LDS share of
// its class's team moves to the top class, where a whole wavefront and its
// LDS budget take it alone.  Scheduling only: results do not depend on it.
template <typename D>
__device__ __forceinline__ int queue_class(const D& d, int nc, int nbi) {
  int c = size_class(nc);
  if (c >= d.heavy_class && c < kSizeClasses - 1) {
    const int ts = team_size(c, d.team_full_class, d.team_max);
    if (ts < kSolveBlock && team_lds_bytes(nbi) > d.team_lds / (kSolveBlock / ts)) c = kSizeClasses - 1;
  }
  return c;
}

// The sweep period of an island whose contacts have one-sweep dependency levels 1..depth and
// whose dynamic bodies see their first and last contact of a sweep at most span levels apart.
// A team lane takes its solves (sweep s, contact of level l) in order of s * period + l: with
// period > span every dependency (inside a sweep, and from a body's last contact of sweep s - 1
// to its first of sweep s) goes to a larger key, so the lanes' orders are consistent and the
// dataflow cannot deadlock; with 2 * period >= depth no more than two sweeps interleave in a
// lane.  Sweeps thereby overlap in the pipeline: the passes approach the dependency depth of
// all sweeps together instead of a sweep-by-sweep walk.
__host__ __device__ inline int sweep_period(int span, int depth) {
  const int p = span + 1, h = (depth + 1) / 2;
  const int r = p > h ? p : h;
  return r < 0xFFFF ? r : 0xFFFF;
}

template <typename T>
struct V4 {
  T x, y, z, w;
};

// Island-body state.  Positions, angles and velocities are FP64 in both
// pipelines: they accumulate many small increments that cancel (weight vs.
// contact impulses), which FP32 would resolve only to ~1e-4 relative.
template <typename T>
struct PV {
  double cx, cy, a, vx, vy, w;
};

// Velocity-constraint point: FP64 lever arms and accumulated impulses, T
// effective masses and bias.  The lever arms stay FP64 in the FP32 pipeline:
// in statically indeterminate piles the load split after the velocity
// iterations moves with the two points' lever arms (through K), and FP32 arms
// alone put the one-step impulse error near 1e-5 of m g dt; rounding the
// masses, effective masses or bias to FP32 moves it by < 1e-6.
template <typename T>
struct VCP {
  double rax, ray, rbx, rby;
  T nm, tm, bias, pad;
  double ni, ti;
};
template <typename T>
struct VC {
  double nx, ny;  // world normal, FP64 for the same reason as the lever arms
  T friction, restitution, ma, mb, Ia, Ib;
  double k11, k12, k22, n11, n12, n21, n22;  // K and K^-1 (n11 = c1x, n21 = c1y, n12 = c2x, n22 = c2y)
  int ia, ib;
  int rka, rkb;  // team path: rank | degree << 16 of the contact in body A's (B's) island sequence
  int npts;
  uint32_t nxt;  // team path: the level-order entry a team size after this contact's (its lane's next)
  VCP<T> p[2];
};
// A point's accumulated normal and tangent impulses, stored together by the team solver.
struct ImpPair {
  double ni, ti;
};
static_assert(offsetof(VCP<float>, ti) == offsetof(VCP<float>, ni) + 8 && sizeof(VCP<float>) % 16 == 0, "VCP");
static_assert(offsetof(VCP<double>, ti) == offsetof(VCP<double>, ni) + 8 && sizeof(VCP<double>) % 16 == 0, "VCP");
template <typename T>
struct PC {
  T lp0x, lp0y, lp1x, lp1y, lnx, lny, lpx, lpy;
  T ra, rb, ma, mb, Ia, Ib, lcax, lcay, lcbx, lcby;
  int type, npts, ia, ib;
  int rka, rkb;         // as in VC
  uint32_t nxt, pad;    // as in VC
  double scx, scy, sa;  // position of the contact's static body (static bodies never move)
};

template <typename T>
struct Dev {
  int nw, nb, nf, ncap;
  double dt, inv_dt;
  int vel_iters, pos_iters;
  int heavy_class;      // islands of this size class and above are solved by a team of lanes
  int team_full_class;  // ... by a whole wavefront from this class up (half one class below, ...)
  int team_lds;         // LDS bytes per wavefront of teams
  int team_max;         // largest team below the top class (8, 16, 32 or 64)
  // worlds
  const int* w_body_off;
  const int* w_fix_off;
  const int* w_con_base;
  const int* w_con_cap;
  int* w_con_used;
  int* w_flags;
  const double* w_grav;
  double* w_inv_dt0;
  int* w_err;
  unsigned long long* stat;  // per-step work counters, summed over steps (S_* below)
  int* w_new;                // contacts created by the world's last pair search (the last ones of its list)
  // contact events (b2g_events): kinds enabled (1 << B2G_EV_*), ring of ev_cap records
  int ev_mask, ev_cap;
  int solve_only;  // b2g_solve_islands: the island search without collide's awake gating
  int* ev_count;
  int* ev_i;
  double* ev_d;
  int* w_move_n;
  int* w_asleep;
  int* w_nisl;
  int2* w_itot;        // per world: bodies and contacts of its islands (the island lists' lengths)
  int* w_ipre;         // exclusive prefixes of w_itot over worlds: bodies [nw + 1], then contacts [nw + 1]
  // bodies
  int* b_type;
  int* b_flags;
  int* b_island;
  double2* b_c;
  double* b_a;
  double2* b_c0;
  double* b_a0;
  double4* b_xf;
  double2* b_lc;
  double* b_sleep;
  double* b_alpha0;  // continuous physics: Sweep.alpha0, and the TOI island index + 1 (0: not in it)
  int* b_tflag;
  V2<double>* b_v;
  double* b_w;
  T* b_im;
  T* b_iI;
  V2<T>* b_f;
  T* b_t;
  T* b_gs;
  T* b_ld;
  T* b_ad;
  // fixtures
  int4* f_info;  // x body (world-local), y shape | nverts << 8 | edge flags << 16, z category | mask << 16, w group
  int2* f_wb;    // world, global body
  double* f_geo;    // shape geometry in FP64 in both precisions: rounding vertices, normals or radii
                    // to FP32 moves near-ties of reference-face choice and clipping (a different
                    // manifold for the whole island), so the narrowphase reads exact inputs
  double* f_geo64;
  double* f_rad64;
  V2<T>* f_mat;
  double4* f_fat;
  int* f_moved;
  int* mv;  // move buffer, per-world segment at w_fix_off[w]
  // contacts (per-world segments at w_con_base[w], capacity w_con_cap[w])
  int2* c_pair;
  int* c_state;
  V2<T>* c_mat;
  V4<T>* c_lnlp;
  V4<T>* c_pts;
  V4<double>* c_imp;  // accumulated impulses: FP64 in both pipelines (warm-start values)
  uint2* c_id;
  double* c_toi;  // continuous physics: the cached time of impact, and toi flag | island flag << 1 |
  int* c_tmeta;   // sub-step count << 8 (contact.py:64-81)
  // islands
  int* i_bodies;     // per-world segment at w_body_off[w]
  int* i_contacts;   // per-world segment at w_con_base[w]
  int2* i_cidx;      // per-world segment at w_con_base[w]
  int* i_isb;        // per-world segment at w_con_base[w]: the world-local start of the contact's
                     // island in the island-body list
  uint2* i_fb;       // per-world segment at w_con_base[w]: the contact's fixtures (x = A | B << 16) and
                     // bodies (y = A | B << 16), world-local, so the solver reads them in one step
  uint2* i_rank;     // per-world segment at w_con_base[w]: the contact's rank in body A's and body B's
                     // contact sequence of the island (x = rank A | rank B << 16) and those bodies'
                     // numbers of island contacts (y = degree A | degree B << 16; 0 for a static side)
  uint32_t* i_order; // per-world segment at w_con_base[w]: each island's contacts ordered by
                     // dependency level within one sweep, as island-local index | level key << 16
                     // (the key is the level plus a per-island offset)
  int4* i_info;      // per-world segment at w_body_off[w]:
                     // (bstart, nb | period << 16, cstart, nc | last_allow << 30)
  int4* q_isl;       // global island queue: (world, bstart, nb | period << 16, cstart); the period
                     // orders a solver lane's sweeps (sweep_period)
  int* q_nc;         // (nc | last_allow << 30)
  int* q_count;
  int4* qs_isl;      // the queue ordered by island size, largest first
  int* qs_nc;
  int* q_cls;        // [kQcls]: per-class counts, cursors, then the wave plan of k_solve
  // solver scratch
  PV<T>* pv;   // indexed like i_bodies
  VC<T>* vc;   // indexed like i_contacts
  PC<T>* pc;
  // worlds too large for the LDS-resident kernels (k_islands_big, k_pairs_big_*);
  // the scratch below is allocated only when the batch has such worlds
  const int* w_big;    // per world: 1 = large path
  const int* big_list; // the large worlds
  const int* w_order;  // worlds by descending body count: the island and pair searches take the longest first
  long long* i_time;   // null, or per queued island: start and end ticks, contacts | bodies << 32, lanes
  int* g_bst;          // per body: state bits as in k_islands
  int* g_wake;         // per body: slot of the contact that woke it (awake gating)
  int* g_off;          // per world nb + 1 at w_body_off[w] + w: CSR offsets
  int* g_cur;          // per body
  int* g_lev;          // per body
  int* g_parent;       // per body: union-find
  int* g_seed;         // per body (by component root)
  // joints (b2g_layout.h joint records), per-world segments at w_joint_off[w]; null without joints
  const int* w_joint_off;
  double* j_d;            // B2G_JOINT_ND doubles per joint
  int* j_i;               // B2G_JOINT_NI ints per joint
  const double4* j_mass;  // 4 per joint: (inverse mass, inverse inertia, mass, 0) of bodies a, b, c, d
  const int* bj_off;      // per body (global) + 1: the body's joint list in bj
  const uint32_t* bj;     // world-local joint | other body << 16 (0xFFFF: none), in Body.joints order
  const int* w_jnc_off;   // per world + 1: its entries in jnc
  const uint32_t* jnc;    // sorted body pairs (lower | higher << 16) that a joint keeps from colliding
  int* j_flag;            // per joint: taken by an island this step
  int* i_joints;          // per-world segment at w_joint_off[w]: the islands' joints (world-local)
  int4* i_jidx;           // indexed like i_joints: island-local indices of the joint's bodies a, b, c, d
                          // (-1 - world-local index for a body outside the island's body list)
  int2* i_jr;             // per island, at its first body (global index): its entries in i_joints
  const int* toi_worlds;  // the worlds with continuous physics (k_toi's grid)
  unsigned char* toi_scr; // per such world: the TOI island's scratch (ToiScratch)
  int* toi_move;          // per world at w_fix_off[w]: the move buffer of a TOI event
  int* g_cntb;         // per body (by component root)
  int* g_cntc;         // per body (by component root)
  int* g_bidx;         // per body: island-local index
  int2* g_adj;         // 2 per contact slot at 2 * w_con_base[w]: (slot, other body)
  int* g_cflag;        // per contact slot
  int* g_stack;        // per world nb + 2 * ccap at w_body_off[w] + 2 * w_con_base[w]
  int4* g_isl;         // per body: island records (root, bstart, cstart, seed)
  int* g_wmisc;        // per world: 16 ints
  int* g_mpos;         // per fixture
  int* g_mlist;        // per fixture
  uint32_t* g_hash;    // per world next_pow2(ccap + 1) at g_hbase[w]
  uint32_t* g_nlist;   // per world next_pow2(ccap) at g_nbase[w]
  const long long* g_hbase;
  const long long* g_nbase;
  // LBVH over each world's fat AABBs (queries, ray casts, the large worlds' tree pair search);
  // per-fixture arrays at w_fix_off[w]
  unsigned long long* t_key;  // leaves in Morton order: 30-bit Morton code << 32 | world-local fixture
  int2* t_child;              // internal node i (0 .. n - 2): children; >= 0 internal, < 0 leaf ~position
  double4* t_box;             // internal node i: the union of its leaves' fat AABBs
  int* t_flag;                // refit state of internal node i
  int* t_root;                // per world: 0 (internal root), ~0 (one leaf) or kTreeEmpty
  unsigned long long* t_sort; // sort scratch of worlds over kTreeLdsKeys fixtures, at t_sbase[w]
  const long long* t_sbase;
};

// ---------------------------------------------------------------------------
// constants (settings.py)
// ---------------------------------------------------------------------------
constexpr double kPi = 3.141592653589793;
constexpr double AABB_EXTENSION = 0.1;
constexpr double AABB_MULTIPLIER = 2.0;
constexpr double LINEAR_SLOP = 0.005;
constexpr double VELOCITY_THRESHOLD = 1.0;
constexpr double MAX_LINEAR_CORRECTION = 0.2;
constexpr double MAX_TRANSLATION = 2.0;
constexpr double MAX_ROTATION = 0.5 * kPi;
constexpr double BAUMGARTE = 0.2;
constexpr double TIME_TO_SLEEP = 0.5;
constexpr double LINEAR_SLEEP_TOL = 0.01;
constexpr double ANGULAR_SLEEP_TOL = 2.0 / 180.0 * kPi;

template <typename T>
__device__ __forceinline__ Geo<double, double> geo_of(const Dev<T>& d, int fg) {
  Geo<double, double> g;
  int4 info = d.f_info[fg];
  g.g = d.f_geo + (size_t)fg * GEO_STRIDE;
  g.type = info.y & 255;
  g.n = (info.y >> 8) & 255;
  g.has0 = (info.y >> 16) & 1;
  g.has3 = (info.y >> 17) & 1;
  return g;
}

__device__ __forceinline__ bool is_static(int type) { return type == BT_STATIC; }


// ---------------------------------------------------------------------------
// contact events: one record per listener call, appended with an atomic and put in the
// reference's call order on the host by (world, phase, key)
// ---------------------------------------------------------------------------
enum : int { EV_BEGIN = 0, EV_END = 1, EV_PRE_SOLVE = 2, EV_POST_SOLVE = 3 };

template <typename T>
__device__ __forceinline__ int ev_open(const Dev<T>& d, int kind) {
  if (!(d.ev_mask & (1 << kind))) return -1;
  const int k = atomicAdd(d.ev_count, 1);
  return k < d.ev_cap ? k : -1;  // the host sees the overflow in the count
}

template <typename T>
__device__ __forceinline__ void ev_manifold(double* o, int* oi, int mtype, int pc, V4<T> lnlp, V4<T> pts,
                                            V4<double> imp, uint2 ids) {
  o[0] = (double)lnlp.x; o[1] = (double)lnlp.y; o[2] = (double)lnlp.z; o[3] = (double)lnlp.w;
  o[4] = (double)pts.x; o[5] = (double)pts.y; o[6] = imp.x; o[7] = imp.y;
  o[8] = (double)pts.z; o[9] = (double)pts.w; o[10] = imp.z; o[11] = imp.w;
  oi[0] = mtype;
  oi[1] = pc;
  oi[2] = (int)ids.x;
  oi[3] = (int)ids.y;
}

// An event of contact slot s of world w: its pair, state flags and current manifold.
template <typename T>
__device__ __forceinline__ void ev_write(const Dev<T>& d, int k, int w, int kind, int phase, int key, int s, int st) {
  int* ei = d.ev_i + (size_t)k * B2G_EV_NI;
  double* ed = d.ev_d + (size_t)k * B2G_EV_ND;
  const int2 pr = d.c_pair[s];
  ei[B2G_EI_WORLD] = w;
  ei[B2G_EI_KIND] = kind;
  ei[B2G_EI_PHASE] = phase;
  ei[B2G_EI_KEY] = key;
  ei[B2G_EI_FA] = pr.x;
  ei[B2G_EI_FB] = pr.y;
  ei[B2G_EI_FLAGS] = ((st & CS_TOUCHING) ? B2G_CF_TOUCHING : 0) | ((st & CS_ENABLED) ? B2G_CF_ENABLED : 0);
  ev_manifold<T>(ed, ei + B2G_EI_MTYPE, (st >> CS_MTYPE_SHIFT) & 3, (st >> CS_PC_SHIFT) & 3, d.c_lnlp[s], d.c_pts[s],
                 d.c_imp[s], d.c_id[s]);
  for (int k = B2G_EI_OLD_MTYPE; k < B2G_EV_NI; ++k) ei[k] = 0;  // the record slot holds an older step's
  for (int k = B2G_ED_OLD; k < B2G_EV_ND; ++k) ed[k] = 0.0;      // record: the callers fill what they use
}

// ---------------------------------------------------------------------------
// Contact.update (contact.py:210-268) of contact slot s of world w, state word st: the manifold
// at the bodies' transforms, warm-start impulses matched by point id, touching, the listener's
// records; returns the new state word (the manifold is written in place).  In the narrow phase
// (toi_seq null) a change of touching is marked CS_WOKE for the island search, which wakes the
// bodies in collide's order, and records are keyed by slot; inside World.solve_toi (toi_seq:
// the world's record counter) the bodies are woken here and records are keyed in call order.
// ---------------------------------------------------------------------------
template <typename T>
__device__ void wake_body(const Dev<T>& d, int gb) {  // Body.awake = True (body.py:800-812)
  const int fl = d.b_flags[gb];
  if (!(fl & BF_AWAKE)) {
    d.b_sleep[gb] = 0.0;
    d.b_flags[gb] = fl | BF_AWAKE;
  }
}

template <typename T>
__device__ int update_contact(const Dev<T>& d, int w, int s, int st, int* toi_seq) {
  const int f0 = d.w_fix_off[w], b0 = d.w_body_off[w];
  const int2 pr = d.c_pair[s];
  const int fa = f0 + pr.x, fb = f0 + pr.y;
  const int ba = b0 + d.f_info[fa].x, bb = b0 + d.f_info[fb].x;
  const int slot = s - d.w_con_base[w];
  // The manifold is evaluated in FP64 from the stored (T) geometry and stored
  // back as T: FP32 arithmetic in the clipping and frame changes would set the
  // FP32 pipeline's one-step error (5e-5 on stacked boxes), FP32 storage does not.
  Man<double> m;
  m.type = (st >> CS_MTYPE_SHIFT) & 3;
  m.pc = (st >> CS_PC_SHIFT) & 3;
  V4<T> lnlp = d.c_lnlp[s], pts = d.c_pts[s];
  const V4<double> imp = d.c_imp[s];
  uint2 ids = d.c_id[s];
  m.ln = v2<double>(lnlp.x, lnlp.y);
  m.lp = v2<double>(lnlp.z, lnlp.w);
  m.p0 = v2<double>(pts.x, pts.y);
  m.p1 = v2<double>(pts.z, pts.w);
  m.id0 = ids.x;
  m.id1 = ids.y;
  const int old_pc = m.pc;
  const uint32_t oid0 = ids.x, oid1 = ids.y;
  const V4<T> old_lnlp = lnlp, old_pts = pts;
  double4 xa = d.b_xf[ba], xb = d.b_xf[bb];
  XF<double> xfa, xfb;  // relative to body A's position
  xfa.p = v2<double>(0, 0);
  xfa.q.c = xa.z;
  xfa.q.s = xa.w;
  xfb.p = v2<double>(xb.x - xa.x, xb.y - xa.y);
  xfb.q.c = xb.z;
  xfb.q.s = xb.w;
  const Geo<double, double> ga = geo_of(d, fa), gb = geo_of(d, fb);
  if (!evaluate(m, ga, xfa, gb, xfb)) {
    atomicOr(&d.w_err[w], WERR_UNSUPPORTED);
    return st;
  }
  // warm start: match ids (the dict keeps the last duplicate)
  double n0 = 0, t0 = 0, n1 = 0, t1 = 0;
  if (m.pc >= 1) {
    if (old_pc >= 1 && oid0 == m.id0) { n0 = imp.x; t0 = imp.y; }
    if (old_pc >= 2 && oid1 == m.id0) { n0 = imp.z; t0 = imp.w; }
  }
  if (m.pc >= 2) {
    if (old_pc >= 1 && oid0 == m.id1) { n1 = imp.x; t1 = imp.y; }
    if (old_pc >= 2 && oid1 == m.id1) { n1 = imp.z; t1 = imp.w; }
  }
  V4<double> nimp;
  nimp.x = m.pc >= 1 ? n0 : imp.x;
  nimp.y = m.pc >= 1 ? t0 : imp.y;
  nimp.z = m.pc >= 2 ? n1 : imp.z;
  nimp.w = m.pc >= 2 ? t1 : imp.w;
  d.c_imp[s] = nimp;
  lnlp.x = (T)m.ln.x; lnlp.y = (T)m.ln.y; lnlp.z = (T)m.lp.x; lnlp.w = (T)m.lp.y;
  pts.x = (T)m.p0.x; pts.y = (T)m.p0.y; pts.z = (T)m.p1.x; pts.w = (T)m.p1.y;
  d.c_lnlp[s] = lnlp;
  d.c_pts[s] = pts;
  ids.x = m.id0;
  ids.y = m.id1;
  d.c_id[s] = ids;
  const bool touching = m.pc > 0;
  const bool was = (st & CS_TOUCHING) != 0;
  int nst = CS_LIVE | CS_ENABLED | (touching ? CS_TOUCHING : 0) | (m.type << CS_MTYPE_SHIFT) | (m.pc << CS_PC_SHIFT) |
            (st & (CS_FILTER | CS_SWAPPED));
  if (toi_seq) {
    if (touching != was) {
      wake_body(d, ba);
      wake_body(d, bb);
    }
  } else {
    nst |= CS_PROCESSED | (touching != was ? CS_WOKE : 0);
  }
  if (d.ev_mask) {  // the listener calls of Contact.update (contact.py:262-268), in that order
    const int ph = toi_seq ? 3 : 0;
    if (!was && touching) {
      const int k = ev_open(d, EV_BEGIN);
      if (k >= 0) ev_write(d, k, w, EV_BEGIN, ph, toi_seq ? (*toi_seq)++ : 4 * slot, s, nst);
    }
    if (was && !touching) {
      const int k = ev_open(d, EV_END);
      if (k >= 0) ev_write(d, k, w, EV_END, ph, toi_seq ? (*toi_seq)++ : 4 * slot + 1, s, nst);
    }
    if (touching) {
      const int k = ev_open(d, EV_PRE_SOLVE);
      if (k >= 0) {
        ev_write(d, k, w, EV_PRE_SOLVE, ph, toi_seq ? (*toi_seq)++ : 4 * slot + 2, s, nst);
        ev_manifold<T>(d.ev_d + (size_t)k * B2G_EV_ND + B2G_ED_OLD, d.ev_i + (size_t)k * B2G_EV_NI + B2G_EI_OLD_MTYPE,
                       (st >> CS_MTYPE_SHIFT) & 3, old_pc, old_lnlp, old_pts, imp, make_uint2(oid0, oid1));
      }
    }
  }
  return nst;
}

// ---------------------------------------------------------------------------
// narrow phase for one contact slot (ContactManager.collide's loop, contact_manager.py:174-223):
// the filter test of a flagged contact, the broad-phase overlap test, then Contact.update;
// returns the new state word (0: destroyed)
// ---------------------------------------------------------------------------
// Body._should_collide's joint test (body.py:920-926): does a joint between the two bodies
// (world-local) keep them from colliding?  A binary search over the world's sorted pairs.
template <typename T>
__device__ __forceinline__ bool joint_blocks(const Dev<T>& d, int w, int la, int lb) {
  if (!d.w_jnc_off) return false;
  int lo = d.w_jnc_off[w], hi = d.w_jnc_off[w + 1];
  const uint32_t key = la < lb ? ((uint32_t)la | (uint32_t)lb << 16) : ((uint32_t)lb | (uint32_t)la << 16);
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    const uint32_t v = d.jnc[mid];
    if (v == key) return true;
    if (v < key) lo = mid + 1;
    else hi = mid;
  }
  return false;
}

template <typename T>
__device__ int narrow_update(const Dev<T>& d, int w, int s, int st) {
  const int f0 = d.w_fix_off[w], b0 = d.w_body_off[w];
  int2 pr = d.c_pair[s];
  const int fa = f0 + pr.x, fb = f0 + pr.y;
  int4 ia = d.f_info[fa], ib = d.f_info[fb];
  const int ba = b0 + ia.x, bb = b0 + ib.x;
  const int slot = s - d.w_con_base[w];
  // a contact destroyed by collide while touching: end_contact (ContactManager.destroy,
  // contact_manager.py:160-172), after the loop's updates
  auto destroyed = [&]() {
    if (st & CS_TOUCHING) {
      const int k = ev_open(d, EV_END);
      if (k >= 0) ev_write(d, k, w, EV_END, 1, slot, s, st);
    }
    return 0;
  };
  if (st & CS_FILTER) {  // refilter (contact_manager.py:200-208)
    bool ok = (d.b_type[ba] == BT_DYNAMIC || d.b_type[bb] == BT_DYNAMIC) && !joint_blocks(d, w, ia.x, ib.x);
    if (ok) {
      if (d.w_flags[w] & WF_CUSTOM_FILTER) ok = !(st & CS_REJECT);  // the host ran the custom filter
      else if (ia.w == ib.w && ia.w != 0) ok = ia.w > 0;  // World.default_contact_filter
      else ok = ((ia.z >> 16) & (ib.z & 0xFFFF)) != 0 && ((ia.z & 0xFFFF) & (ib.z >> 16)) != 0;
    }
    if (!ok) return destroyed();
    st &= ~(CS_FILTER | CS_REJECT);
  }
  double4 A = d.f_fat[fa], B = d.f_fat[fb];
  if (!aabb_overlap(A.x, A.y, A.z, A.w, B.x, B.y, B.z, B.w)) return destroyed();
  return update_contact(d, w, s, st, nullptr);
}

template <typename T, int B>
__global__ void __launch_bounds__(B) k_collide(Dev<T> d, int chunks) {
  constexpr int kBlock = B;  // contact slots per workgroup (B2G_PARAM_COLLIDE_BLOCK)
  // Workgroups are dealt to the 8 XCDs round-robin; all chunks of world w get
  // workgroup ids = w (mod 8), so a world's fixtures and bodies are fetched
  // into one XCD's L2 only.
  // The chunk's contacts that need a manifold are then regrouped by shape pair (a counting
  // sort in LDS), so that a wavefront runs one collider instead of all of them in turn.
  __shared__ int count[16], base[16], total_s, live_s;
  __shared__ uint16_t sorted[kBlock];
  const int b = blockIdx.x, xcd = b & 7, k = b >> 3;
  const int w = (k / chunks) * 8 + xcd, chunk = k % chunks;
  if (w >= d.nw) return;  // the whole workgroup
  const int used = d.w_con_used[w];
  if (chunk * kBlock >= used) return;  // the whole workgroup: past the world's used slots
  const int tid = threadIdx.x;
  if (tid < 16) count[tid] = 0;
  if (tid == 0) live_s = 0;
  __syncthreads();
  const int sl = chunk * kBlock + tid;
  const int f0 = d.w_fix_off[w], b0 = d.w_body_off[w], cbase = d.w_con_base[w];
  int key = -1, rank = 0;
  if (sl < used) {
    const int s = cbase + sl;
    int st = d.c_state[s];
    if (st & CS_LIVE) {
      atomicAdd(&live_s, 1);
      const int2 pr = d.c_pair[s];
      const int4 ia = d.f_info[f0 + pr.x], ib = d.f_info[f0 + pr.y];
      const int ba = b0 + ia.x, bb = b0 + ib.x;
      const bool act_a = (d.b_flags[ba] & BF_AWAKE) && !is_static(d.b_type[ba]);
      const bool act_b = (d.b_flags[bb] & BF_AWAKE) && !is_static(d.b_type[bb]);
      if (act_a || act_b) {
        key = ((ia.y & 3) << 2) | (ib.y & 3);  // shape types of A and B
        rank = atomicAdd(&count[key], 1);
      } else {
        d.c_state[s] = st & ~(CS_PROCESSED | CS_WOKE);
      }
    }
  }
  __syncthreads();
  if (tid == 0) {
    int run = 0;
    for (int q = 0; q < 16; ++q) {
      base[q] = run;
      run += count[q];
    }
    total_s = run;
    atomicAdd(&d.stat[S_LIVE], (unsigned long long)live_s);
    atomicAdd(&d.stat[S_EVAL], (unsigned long long)run);
  }
  __syncthreads();
  const int total = total_s;
  if (key >= 0) sorted[base[key] + rank] = (uint16_t)tid;
  __syncthreads();
  if (tid < total) {
    const int s = cbase + chunk * kBlock + sorted[tid];
    const int st = d.c_state[s] & ~(CS_PROCESSED | CS_WOKE);
    d.c_state[s] = narrow_update(d, w, s, st);
  }
}

// ---------------------------------------------------------------------------
// per-world island search
// ---------------------------------------------------------------------------
// The 32-bit LDS address of a pointer into the workgroup's LDS (for LDS reads written in inline
// assembly).
template <typename P>
__device__ __forceinline__ uint32_t lds_addr(P* p) {
  return (uint32_t)(uintptr_t)((__attribute__((address_space(3))) P*)p);
}
struct IslandSmem {
  int* off;        // nb + 1
  int* cur;        // nb
  uint32_t* adj;    // adjacency entries: contact slot | other body << 16
  uint8_t* bst;    // nb: bit0 static, bit1 awake, bit2 allow_sleep, bit3 active, bit4 flag, bit5 woken
  uint16_t* bidx;  // nb
  uint16_t* stack;  // nb
  uint16_t* slist;  // nb
  uint16_t* lev;    // nb: dependency level of the body's last contact in the island list
  uint16_t* ibs;    // nb: start of the body's island in the island-body list
  uint32_t* cflag;  // ccap / 32
  int* scan;       // kBlock
  int* misc;       // 8
};

__device__ __forceinline__ size_t align16(size_t x) { return (x + 15) & ~(size_t)15; }

// Touching contacts of a world are a planar graph over its fixtures (at most
// 3 per fixture on average); the island search keeps room for 4 per fixture
// and reports a capacity error beyond that.  The adjacency arrays also hold
// the awake-gating candidates, kGateChunk slots at a time.
constexpr int kGateChunk = 1024;
__host__ __device__ inline int touch_cap(int nf, int ccap) {
  const int t = 4 * nf + 64;
  return t < ccap ? t : ccap;
}
__host__ __device__ inline int adj_entries(int nf, int ccap) {
  const int e = 2 * touch_cap(nf, ccap);
  return e > 2 * kGateChunk ? e : 2 * kGateChunk;
}

__host__ __device__ inline size_t island_smem_bytes(int nb, int nf, int ccap) {
  size_t s = 0;
  s += ((size_t)(nb + 1) * 4 + 15) & ~(size_t)15;
  s += ((size_t)nb * 4 + 15) & ~(size_t)15;
  s += ((size_t)adj_entries(nf, ccap) * 2 + 15) & ~(size_t)15;
  s += ((size_t)adj_entries(nf, ccap) * 2 + 15) & ~(size_t)15;
  s += ((size_t)nb + 15) & ~(size_t)15;
  s += ((size_t)nb * 2 + 15) & ~(size_t)15;
  s += ((size_t)nb * 2 + 15) & ~(size_t)15;
  s += ((size_t)nb * 2 + 15) & ~(size_t)15;
  s += ((size_t)nb * 2 + 15) & ~(size_t)15;  // lev
  s += ((size_t)nb * 2 + 15) & ~(size_t)15;  // ibs
  s += ((size_t)(ccap + 31) / 32 * 4 + 15) & ~(size_t)15;
  s += (size_t)kBlock * 4;
  s += 64 * 4;  // misc: 16 scalars, 16 class counts, spare
  return s;
}

// Exclusive scan of a[0..n) in place by one workgroup; returns the total.
template <int B = kBlock>
__device__ int block_exclusive_scan(int* a, int n, int* scratch) {
  constexpr int kBlock = B;
  const int tid = threadIdx.x;
  const int per = (n + kBlock - 1) / kBlock;
  const int beg = tid * per, end = min(n, beg + per);
  int sum = 0;
  for (int i = beg; i < end; ++i) sum += a[i];
  scratch[tid] = sum;
  __syncthreads();
  for (int off = 1; off < kBlock; off <<= 1) {
    int v = tid >= off ? scratch[tid - off] : 0;
    __syncthreads();
    scratch[tid] += v;
    __syncthreads();
  }
  int run = scratch[tid] - sum;
  for (int i = beg; i < end; ++i) {
    int v = a[i];
    a[i] = run;
    run += v;
  }
  __syncthreads();
  int total = scratch[kBlock - 1];
  __syncthreads();
  return total;
}

// B threads per workgroup: 256, or 128 for worlds whose LDS footprint lets more than five
// workgroups share a CU (the one-lane walk dominates the kernel, so more worlds in flight
// per CU outweigh the slower wide phases).
template <typename T, int B>
__global__ void __launch_bounds__(B) __attribute__((amdgpu_waves_per_eu(5))) k_islands(Dev<T> d, int w_first) {
  constexpr int kBlock = B;
  extern __shared__ __align__(16) unsigned char smem[];
  const int w = d.w_order[w_first + blockIdx.x], tid = threadIdx.x;
  if (d.w_big && d.w_big[w]) return;  // k_islands_big
  const int b0 = d.w_body_off[w], nb = d.w_body_off[w + 1] - b0;
  const int f0 = d.w_fix_off[w], nf = d.w_fix_off[w + 1] - f0;
  const int cb = d.w_con_base[w], ccap = d.w_con_cap[w];
  const int used = d.w_con_used[w];
  const int nadj = adj_entries(nf, ccap);
  IslandSmem L;
  {
    unsigned char* p = smem;
    L.off = (int*)p; p += ((size_t)(nb + 1) * 4 + 15) & ~(size_t)15;
    L.cur = (int*)p; p += ((size_t)nb * 4 + 15) & ~(size_t)15;
    L.adj = (uint32_t*)p; p += ((size_t)nadj * 4 + 15) & ~(size_t)15;
    L.bst = (uint8_t*)p; p += ((size_t)nb + 15) & ~(size_t)15;
    L.bidx = (uint16_t*)p; p += ((size_t)nb * 2 + 15) & ~(size_t)15;
    L.stack = (uint16_t*)p; p += ((size_t)nb * 2 + 15) & ~(size_t)15;
    L.slist = (uint16_t*)p; p += ((size_t)nb * 2 + 15) & ~(size_t)15;
    L.lev = (uint16_t*)p; p += ((size_t)nb * 2 + 15) & ~(size_t)15;
    L.ibs = (uint16_t*)p; p += ((size_t)nb * 2 + 15) & ~(size_t)15;
    L.cflag = (uint32_t*)p; p += ((size_t)(ccap + 31) / 32 * 4 + 15) & ~(size_t)15;
    L.scan = (int*)p; p += (size_t)kBlock * 4;
    L.misc = (int*)p;
  }
  for (int i = tid; i < nb; i += kBlock) {
    int fl = d.b_flags[b0 + i];
    int ty = d.b_type[b0 + i];
    L.bst[i] = (uint8_t)((ty == BT_STATIC ? 1 : 0) | ((fl & BF_AWAKE) ? 2 : 0) | ((fl & BF_ALLOW_SLEEP) ? 4 : 0) |
                         ((fl & BF_ACTIVE) ? 8 : 0));
    L.off[i] = 0;
  }
  for (int i = tid; i < (ccap + 31) / 32; i += kBlock) L.cflag[i] = 0;
  if (tid == 0) L.misc[4] = 0;
  __syncthreads();

  // --- rest of collide's awake gating (contact_manager.py:191-197), in order ---
  // Only live contacts that k_collide skipped (both bodies asleep) or whose
  // touching flag changed (which wakes bodies) can matter.  They are first
  // compacted in slot order, with their bodies, into LDS (the adjacency arrays
  // are free until the graph is built); one lane then walks them in order.
  if (d.w_asleep[w] > 0 && !d.solve_only) {
    uint16_t* cs = (uint16_t*)L.adj;  // the adjacency area holds 4 * kGateChunk halves at least
    uint16_t* cla = cs + kGateChunk;
    uint16_t* clb = cla + kGateChunk;
    uint16_t* cfl = clb + kGateChunk;  // bit 0 evaluated this step, bit 1 touching changed
    int* wcount = L.scan;
    const int lane = tid & 63, wv = tid >> 6;
    for (int chunk = 0; chunk < used; chunk += kGateChunk) {
      const int cend = min(used, chunk + kGateChunk);
      for (int base = chunk; base < cend; base += kBlock) {
        const int sl = base + tid;
        int st = 0;
        bool cand = false;
        if (sl < cend) {
          st = d.c_state[cb + sl];
          cand = (st & CS_LIVE) && (!(st & CS_PROCESSED) || (st & CS_WOKE));
        }
        const unsigned long long m = __ballot(cand);
        if (lane == 0) wcount[wv] = __popcll(m);
        __syncthreads();
        int off = L.misc[4];
        for (int j = 0; j < wv; ++j) off += wcount[j];
        if (cand) {
          const int idx = off + __popcll(m & ((1ull << lane) - 1));
          const int2 pr = d.c_pair[cb + sl];
          cs[idx] = (uint16_t)sl;
          cla[idx] = (uint16_t)d.f_info[f0 + pr.x].x;
          clb[idx] = (uint16_t)d.f_info[f0 + pr.y].x;
          cfl[idx] = (uint16_t)(((st & CS_PROCESSED) ? 1 : 0) | ((st & CS_WOKE) ? 2 : 0));
        }
        __syncthreads();
        if (tid == 0) {
          int add = 0;
          for (int j = 0; j < kBlock / 64; ++j) add += wcount[j];
          L.misc[4] += add;
        }
        __syncthreads();
      }
      if (tid == 0) {
        const int ncand = L.misc[4];
        for (int i = 0; i < ncand; ++i) {
          const int la = cla[i], lb = clb[i], fl = cfl[i];
          bool woke = (fl & 2) != 0;
          if (!(fl & 1)) {
            const bool act = ((L.bst[la] & 2) && !(L.bst[la] & 1)) || ((L.bst[lb] & 2) && !(L.bst[lb] & 1));
            if (!act) continue;
            const int s = cb + cs[i];
            const int st = narrow_update(d, w, s, d.c_state[s]);
            d.c_state[s] = st;
            if (!(st & CS_LIVE)) continue;
            woke = (st & CS_WOKE) != 0;
          }
          if (woke) {
            if (!(L.bst[la] & 2)) L.bst[la] |= 2 | 32;
            if (!(L.bst[lb] & 2)) L.bst[lb] |= 2 | 32;
          }
        }
        L.misc[4] = 0;
      }
      __syncthreads();
    }
  }
  __syncthreads();

  // --- touching-contact graph in ContactManager order (body._contacts order) ---
  for (int sl = tid; sl < used; sl += kBlock) {
    int st = d.c_state[cb + sl];
    if ((st & (CS_LIVE | CS_TOUCHING | CS_ENABLED)) == (CS_LIVE | CS_TOUCHING | CS_ENABLED)) {
      int2 pr = d.c_pair[cb + sl];
      atomicAdd(&L.off[d.f_info[f0 + pr.x].x], 1);
      atomicAdd(&L.off[d.f_info[f0 + pr.y].x], 1);
    }
  }
  __syncthreads();
  int total = block_exclusive_scan<B>(L.off, nb, L.scan);
  if (total > nadj) {  // more touching contacts than the island search has room for
    if (tid == 0) {
      atomicOr(&d.w_err[w], WERR_CAPACITY);
      d.w_nisl[w] = 0;
      d.w_itot[w] = make_int2(0, 0);
    }
    return;
  }
  if (tid == 0) L.off[nb] = total;
  for (int i = tid; i < nb; i += kBlock) L.cur[i] = L.off[i];
  __syncthreads();
  for (int sl = tid; sl < used; sl += kBlock) {
    int st = d.c_state[cb + sl];
    if ((st & (CS_LIVE | CS_TOUCHING | CS_ENABLED)) == (CS_LIVE | CS_TOUCHING | CS_ENABLED)) {
      int2 pr = d.c_pair[cb + sl];
      int la = d.f_info[f0 + pr.x].x, lb = d.f_info[f0 + pr.y].x;
      int pa = atomicAdd(&L.cur[la], 1);
      L.adj[pa] = (uint32_t)sl | (uint32_t)lb << 16;
      int pb = atomicAdd(&L.cur[lb], 1);
      L.adj[pb] = (uint32_t)sl | (uint32_t)la << 16;
    }
  }
  __syncthreads();
  const int j0 = d.w_joint_off ? d.w_joint_off[w] : 0;
  const int nj = d.w_joint_off ? d.w_joint_off[w + 1] - j0 : 0;
  for (int k = tid; k < nj; k += kBlock) d.j_flag[j0 + k] = 0;
  for (int i = tid; i < nb; i += kBlock) {  // restore ascending slot order per body
    L.cur[i] = 0;  // from here on: contacts of the body seen by the island search
    L.lev[i] = 0;
    int beg = L.off[i], end = L.off[i + 1];
    for (int k = beg + 1; k < end; ++k) {
      const uint32_t e = L.adj[k];
      int j = k - 1;
      while (j >= beg && (L.adj[j] & 0xFFFFu) > (e & 0xFFFFu)) {
        L.adj[j + 1] = L.adj[j];
        --j;
      }
      L.adj[j + 1] = e;
    }
  }
  __syncthreads();

  // --- depth-first island search (world.py:494-580), one lane ---
  int* isl_b = d.i_bodies + b0;
  int* isl_c = d.i_contacts + cb;
  int4* info = d.i_info + b0;
  if (tid == 0) {
    int nbi_tot = 0, nci_tot = 0, nisl = 0, njt = 0;
    for (int seed = 0; seed < nb; ++seed) {
      uint8_t sb = L.bst[seed];
      if (sb & 16) continue;            // island flag
      if (!(sb & 2) || !(sb & 8)) continue;  // not awake or not active
      if (sb & 1) continue;             // static
      int sp = 0, ns = 0;
      L.stack[sp++] = (uint16_t)seed;
      L.bst[seed] = sb | 16;
      const int bstart = nbi_tot, cstart = nci_tot, jstart = njt;
      int last_allow = 0, span = 0, depth = 0;
      while (sp > 0) {
        const int b = L.stack[--sp];
        uint8_t bs = L.bst[b];
        last_allow = (bs >> 2) & 1;
        if (!(bs & 2)) bs |= 2 | 32;  // body.awake = True
        L.bst[b] = bs;
        if (bs & 1) {  // static bodies join but do not propagate
          L.slist[ns++] = (uint16_t)b;
          continue;
        }
        L.bidx[b] = (uint16_t)nbi_tot;
        L.ibs[b] = (uint16_t)bstart;
        isl_b[nbi_tot++] = b;
        // the expanding body's counters stay in registers; cur = rank | first level << 16
        int cur_b = L.cur[b], lev_b = L.lev[b];
        // The next entry is fetched one step ahead, and an entry's contact flag and its other
        // body's state are read together (also for entries that turn out to be taken), so an
        // entry costs about one LDS round trip.
        const int kend = L.off[b + 1];
        uint32_t e = L.off[b] < kend ? L.adj[L.off[b]] : 0u;
        for (int k = L.off[b]; k < kend; ++k) {
          const int s = (int)(e & 0xFFFFu), o = (int)(e >> 16);
          const uint32_t bit = 1u << (s & 31);
          // the entry's contact flag word, the other body's state, rank counter and level, and
          // the body's next entry: five LDS reads issued together and waited for once
          uint32_t fw, so32, co32, lo32;
          asm volatile(
              "ds_read_b32 %0, %5\n\t"
              "ds_read_u8 %1, %6\n\t"
              "ds_read_b32 %2, %7\n\t"
              "ds_read_u16 %3, %8\n\t"
              "ds_read_b32 %4, %9\n\t"
              "s_waitcnt lgkmcnt(0)"
              : "=&v"(fw), "=&v"(so32), "=&v"(co32), "=&v"(lo32), "=&v"(e)
              : "v"(lds_addr(&L.cflag[s >> 5])), "v"(lds_addr(&L.bst[o])), "v"(lds_addr(&L.cur[o])),
                "v"(lds_addr(&L.lev[o])), "v"(lds_addr(&L.adj[k + 1 < kend ? k + 1 : k]))
              : "memory");
          const uint8_t so = (uint8_t)so32;
          const int co = (int)co32;
          const int lo_read = (int)lo32;
          if (fw & bit) continue;
          L.cflag[s >> 5] = fw | bit;
          const bool o_dyn = !(so & 1);
          // level: 1 + the latest level among the contact's dynamic bodies (one sweep)
          const int lo = o_dyn ? lo_read : 0;
          const int lvl = 1 + (lev_b > lo ? lev_b : lo);
          const uint32_t rank_b = (uint32_t)(cur_b & 0xFFFF);
          const uint32_t rank_o = o_dyn ? (uint32_t)(co & 0xFFFF) : 0u;
          cur_b += 1 + (rank_b == 0 ? lvl << 16 : 0);  // a first contact records its level
          if (o_dyn) L.cur[o] = co + 1 + (rank_o == 0 ? lvl << 16 : 0);
          lev_b = lvl;
          if (o_dyn) L.lev[o] = (uint16_t)lvl;
          const uint32_t key = (uint32_t)(cstart + lvl - 1);  // < cstart + island contacts
          d.i_rank[cb + nci_tot] = make_uint2((uint32_t)b, rank_b | rank_o << 16);
          d.i_order[cb + nci_tot] = (uint32_t)(nci_tot - cstart) | key << 16;  // sorted below
          isl_c[nci_tot++] = s;
          if (so & 16) continue;
          L.stack[sp++] = (uint16_t)o;
          L.bst[o] = so | 16;
        }
        L.cur[b] = cur_b;
        L.lev[b] = (uint16_t)lev_b;
        // b has no contacts left to take: its levels run from its first contact's to lev_b
        if (cur_b & 0xFFFF) {
          span = max(span, lev_b - (int)((uint32_t)cur_b >> 16));
          depth = max(depth, lev_b);
        }
        // the body's joints (world.py:548-572), in Body.joints order
        if (nj > 0) {
          const int gb = b0 + b;
          for (int k = d.bj_off[gb]; k < d.bj_off[gb + 1]; ++k) {
            const uint32_t e = d.bj[k];
            const int jl = (int)(e & 0xFFFFu), o = (int)(e >> 16);
            if (d.j_flag[j0 + jl]) continue;
            if (o == 0xFFFF) {  // a joint on one body (mouse)
              d.i_joints[j0 + njt++] = jl;
              d.j_flag[j0 + jl] = 1;
              continue;
            }
            const uint8_t so = L.bst[o];
            if (!(so & 8)) continue;  // joints to inactive bodies are not simulated
            d.i_joints[j0 + njt++] = jl;
            d.j_flag[j0 + jl] = 1;
            if (so & 16) continue;
            L.stack[sp++] = (uint16_t)o;
            L.bst[o] = so | 16;
          }
        }
      }
      for (int k = 0; k < ns; ++k) L.bst[L.slist[k]] &= (uint8_t)~16;  // statics may join other islands
      info[nisl++] = make_int4(bstart, (nbi_tot - bstart) | sweep_period(span, depth) << 16, cstart,
                               (nci_tot - cstart) | (last_allow << 30));
      if (nj > 0) d.i_jr[b0 + bstart] = make_int2(j0 + jstart, njt - jstart);
    }
    L.misc[0] = nisl;
    L.misc[1] = nbi_tot;
    L.misc[2] = nci_tot;
    L.misc[5] = njt;
    L.misc[3] = nisl > 0 ? atomicAdd(d.q_count, nisl) : 0;
  }
  __syncthreads();
  const int nisl = L.misc[0], nci_tot = L.misc[2], qbase = L.misc[3];
  // the island joints' bodies as island-local indices (a body outside the island's body list,
  // a static one, as -1 - world-local index)
  for (int k = tid; k < L.misc[5]; k += kBlock) {
    const int* ji = d.j_i + (size_t)(j0 + d.i_joints[j0 + k]) * B2G_JOINT_NI;
    const int ba = ji[B2G_JI_BODY_A], bb = ji[B2G_JI_BODY_B];
    const int home = !(L.bst[bb] & 1) ? bb : ba;  // a body of the joint's island
    const int ist = L.ibs[home];
    const bool gear = ji[B2G_JI_TYPE] == B2G_JT_GEAR;  // only gear joints have bodies c, d (checked at upload)
    const int slot[4] = {ba < 0 ? bb : ba, bb, gear ? ji[B2G_JI_BODY_C] : bb, gear ? ji[B2G_JI_BODY_D] : bb};
    int idx[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int x = slot[q];
      const uint8_t bs = L.bst[x];
      idx[q] = (!(bs & 1) && (bs & 16) && L.ibs[x] == ist) ? (int)L.bidx[x] - ist : -1 - x;
    }
    d.i_jidx[j0 + k] = make_int4(idx[0], idx[1], idx[2], idx[3]);
  }
  int* ccount = L.misc + 16;  // islands per size class, then one global add per class
  if (tid < kSizeClasses) ccount[tid] = 0;
  __syncthreads();
  for (int k = tid; k < nisl; k += kBlock) {
    int4 in = info[k];
    d.q_isl[qbase + k] = make_int4(w, b0 + in.x, in.y, cb + in.z);
    d.q_nc[qbase + k] = in.w;
    atomicAdd(&ccount[queue_class(d, in.w & 0x3FFFFFFF, in.y & 0xFFFF)], 1);
  }
  __syncthreads();
  if (tid < kSizeClasses && ccount[tid] > 0) atomicAdd(&d.q_cls[tid], ccount[tid]);
  for (int k = tid; k < nci_tot; k += kBlock) {
    const int s = isl_c[k];
    int2 pr = d.c_pair[cb + s];
    const int la = d.f_info[f0 + pr.x].x, lb = d.f_info[f0 + pr.y].x;
    int ia = (L.bst[la] & 1) ? -1 - la : (int)L.bidx[la];
    int ib = (L.bst[lb] & 1) ? -1 - lb : (int)L.bidx[lb];
    d.i_cidx[cb + k] = make_int2(ia, ib);
    d.i_fb[cb + k] = make_uint2((uint32_t)pr.x | (uint32_t)pr.y << 16, (uint32_t)la | (uint32_t)lb << 16);
    d.i_isb[cb + k] = (int)L.ibs[ia >= 0 ? la : lb];
    // (expanding body, its rank | the other body's rank << 16) -> ranks and degrees of (A, B);
    // a dynamic body's rank counter ends at its number of island contacts
    const uint2 r = d.i_rank[cb + k];
    const uint32_t lo = r.y & 0xFFFFu, hi = r.y >> 16;
    const uint32_t dga = ia >= 0 ? (uint32_t)L.cur[la] & 0xFFFFu : 0u;
    const uint32_t dgb = ib >= 0 ? (uint32_t)L.cur[lb] & 0xFFFFu : 0u;
    d.i_rank[cb + k] = make_uint2((int)r.x == la ? (lo | hi << 16) : (hi | lo << 16), dga | dgb << 16);
  }
  // each island's contacts by level: a counting sort of the entries (island-local index |
  // key << 16, key = island start + level - 1, so keys of one island stay inside its contact
  // range) through LDS; the adjacency arrays are free now (they hold 2 entries per contact)
  int* hist = (int*)L.adj;
  uint32_t* ent = L.adj + nci_tot;
  for (int k = tid; k < nci_tot; k += kBlock) {
    hist[k] = 0;
    ent[k] = d.i_order[cb + k];
  }
  __syncthreads();
  for (int k = tid; k < nci_tot; k += kBlock) atomicAdd(&hist[ent[k] >> 16], 1);
  __syncthreads();
  block_exclusive_scan<B>(hist, nci_tot, L.scan);
  for (int k = tid; k < nci_tot; k += kBlock) {
    const uint32_t e = ent[k];
    d.i_order[cb + atomicAdd(&hist[e >> 16], 1)] = e;
  }
  for (int i = tid; i < nb; i += kBlock) {
    uint8_t bs = L.bst[i];
    int fl = d.b_flags[b0 + i];
    if (bs & 32) {  // woken: Body.awake setter (body.py:800-812)
      fl |= BF_AWAKE;
      d.b_sleep[b0 + i] = 0.0;
      d.b_flags[b0 + i] = fl;
    }
    d.b_island[b0 + i] = ((bs & 16) && !(bs & 1)) ? 1 : -1;
  }
  if (tid == 0) {
    d.w_nisl[w] = nisl;
    d.w_itot[w] = make_int2(L.misc[1], nci_tot);
  }
}

// ---------------------------------------------------------------------------
// Island.solve, one lane per island (island.py:65-179, contact.py:375-864)
//
// In the FP32 pipeline the constraint geometry (world manifold, lever arms,
// normals, effective masses) and the position solver's geometry are FP32;
// positions, angles, velocities, accumulated impulses and the 2x2 block LCP
// are FP64.
// ---------------------------------------------------------------------------
using V2d = V2<double>;

template <typename T>
__device__ __forceinline__ V2d dv2(T x, T y) {
  return v2<double>((double)x, (double)y);
}

template <typename T>
__device__ __forceinline__ void load_pos(const Dev<T>& d, int pvb, int b0, int i, double& cx, double& cy, double& a) {
  if (i >= 0) {
    const PV<T>& p = d.pv[pvb + i];
    cx = p.cx;
    cy = p.cy;
    a = p.a;
  } else {
    double2 c = d.b_c[b0 + (-1 - i)];
    cx = c.x;
    cy = c.y;
    a = d.b_a[b0 + (-1 - i)];
  }
}
template <typename T>
__device__ __forceinline__ void load_vel(const Dev<T>& d, int pvb, int b0, int i, V2d& v, double& w) {
  if (i >= 0) {
    const PV<T>& p = d.pv[pvb + i];
    v = v2<double>(p.vx, p.vy);
    w = p.w;
  } else {  // static bodies never change
    v = d.b_v[b0 + (-1 - i)];
    w = d.b_w[b0 + (-1 - i)];
  }
}
template <typename T>
__device__ __forceinline__ void store_vel(const Dev<T>& d, int pvb, int i, V2d v, double w) {
  if (i >= 0) {
    PV<T>& p = d.pv[pvb + i];
    p.vx = v.x;
    p.vy = v.y;
    p.w = w;
  }
}
template <typename T>
__device__ __forceinline__ void store_pos(const Dev<T>& d, int pvb, int i, double cx, double cy, double a) {
  if (i >= 0) {
    PV<T>& p = d.pv[pvb + i];
    p.cx = cx;
    p.cy = cy;
    p.a = a;
  }
}

// World manifold (contact_util.py:195-245) in a frame centred on body A's centre.
template <typename T>
__device__ V2<T> world_manifold(int type, int pc, V2<T> ln, V2<T> lp, V2<T> p0, V2<T> p1, XF<T> xfa, T ra,
                                XF<T> xfb, T rb, V2<T>& w0, V2<T>& w1) {
  V2<T> normal = v2<T>(0, 0);
  if (pc == 0) return normal;
  if (type == M_CIRCLES) {
    normal = v2<T>(1, 0);
    V2<T> pa = xmul(xfa, lp), pb = xmul(xfb, p0);
    V2<T> dd = pa - pb;
    if (dot(dd, dd) > NormEps<T>::value * NormEps<T>::value) normal = normalized(pb - pa);
    V2<T> ca = pa + scale(ra, normal);
    V2<T> cbp = pb - scale(rb, normal);
    w0 = scale((T)0.5, ca + cbp);
  } else if (type == M_FACE_A) {
    normal = rmul(xfa.q, ln);
    V2<T> plane = xmul(xfa, lp);
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      if (i >= pc) break;
      V2<T> clip = xmul(xfb, i == 0 ? p0 : p1);
      V2<T> ca = clip + scale(ra - dot(clip - plane, normal), normal);
      V2<T> cbp = clip - scale(rb, normal);
      V2<T> r = scale((T)0.5, ca + cbp);
      if (i == 0) w0 = r; else w1 = r;
    }
  } else {
    normal = rmul(xfb.q, ln);
    V2<T> plane = xmul(xfb, lp);
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      if (i >= pc) break;
      V2<T> clip = xmul(xfa, i == 0 ? p0 : p1);
      V2<T> cbp = clip + scale(rb - dot(clip - plane, normal), normal);
      V2<T> ca = clip - scale(ra, normal);
      V2<T> r = scale((T)0.5, ca + cbp);
      if (i == 0) w0 = r; else w1 = r;
    }
    normal = -normal;
  }
  return normal;
}

// The FP32 pipeline evaluates the velocity constraints with fused multiply-adds (one rounding
// per product-and-sum instead of two, and about 40 % fewer FP64 instructions per solve); its
// results are held to the FP64 oracle within the FP32 tolerance.  The FP64 pipeline keeps the
// reference's separate products and sums, operation by operation, and so stays exact.
template <typename T>
struct FusedSolve {
  static constexpr bool value = false;
};
template <>
struct FusedSolve<float> {
  static constexpr bool value = true;
};
__device__ __forceinline__ double fmad(double a, double b, double c) { return __builtin_fma(a, b, c); }

// relative velocity of the contact point (dv = vb + wb x rb - va - wa x ra), fused
__device__ __forceinline__ void rel_vel_fused(V2d va, double wa, V2d vb, double wb, double rax, double ray,
                                              double rbx, double rby, double& dvx, double& dvy) {
  dvx = fmad(-wb, rby, vb.x) - fmad(-wa, ray, va.x);
  dvy = fmad(wb, rbx, vb.y) - fmad(wa, rax, va.y);
}
// applies impulse P at lever arms ra, rb (va -= ma P, wa -= Ia ra x P, vb += mb P, wb += Ib rb x P), fused
__device__ __forceinline__ void apply_fused(double ma, double mb, double Ia, double Ib, double rax, double ray,
                                            double rbx, double rby, double Px, double Py, V2d& va, double& wa,
                                            V2d& vb, double& wb) {
  va.x = fmad(-ma, Px, va.x);
  va.y = fmad(-ma, Py, va.y);
  wa = fmad(-Ia, fmad(rax, Py, -ray * Px), wa);
  vb.x = fmad(mb, Px, vb.x);
  vb.y = fmad(mb, Py, vb.y);
  wb = fmad(Ib, fmad(rbx, Py, -rby * Px), wb);
}

template <typename T>
__device__ __forceinline__ void apply_block(VC<T>& vc, V2d x, V2d a, V2d normal, V2d& va, double& wa, V2d& vb,
                                            double& wb) {
  V2d dd = x - a;
  V2d p1 = scale(dd.x, normal), p2 = scale(dd.y, normal);
  V2d ra1 = dv2(vc.p[0].rax, vc.p[0].ray), ra2 = dv2(vc.p[1].rax, vc.p[1].ray);
  V2d rb1 = dv2(vc.p[0].rbx, vc.p[0].rby), rb2 = dv2(vc.p[1].rbx, vc.p[1].rby);
  va = va - scale((double)vc.ma, p1 + p2);
  wa -= (double)vc.Ia * (cross(ra1, p1) + cross(ra2, p2));
  vb = vb + scale((double)vc.mb, p1 + p2);
  wb += (double)vc.Ib * (cross(rb1, p1) + cross(rb2, p2));
  vc.p[0].ni = x.x;
  vc.p[1].ni = x.y;
}

// One contact of solve_velocity_constraints (contact.py:527-799).
template <typename T>
__device__ __forceinline__ void tangent_point(const VC<T>& vc, VCP<T>& p, V2d tangent, V2d& va, double& wa, V2d& vb,
                                              double& wb) {
  const double ma = vc.ma, mb = vc.mb, Ia = vc.Ia, Ib = vc.Ib;
  if constexpr (FusedSolve<T>::value) {
    const double rax = p.rax, ray = p.ray, rbx = p.rbx, rby = p.rby;
    double dvx, dvy;
    rel_vel_fused(va, wa, vb, wb, rax, ray, rbx, rby, dvx, dvy);
    const double vt = fmad(dvx, tangent.x, dvy * tangent.y);
    double lambda = (double)p.tm * (-vt);
    const double maxf = (double)vc.friction * p.ni;
    const double newi = pclamp(p.ti + lambda, -maxf, maxf);
    lambda = newi - p.ti;
    apply_fused(ma, mb, Ia, Ib, rax, ray, rbx, rby, lambda * tangent.x, lambda * tangent.y, va, wa, vb, wb);
    p.ti = newi;
  } else {
    V2d ra = dv2(p.rax, p.ray), rb = dv2(p.rbx, p.rby);
    V2d dv = vb + cross_sv(wb, rb) - va - cross_sv(wa, ra);
    double vt = dot(dv, tangent);
    double lambda = (double)p.tm * (-vt);
    double maxf = (double)vc.friction * p.ni;
    double newi = pclamp(p.ti + lambda, -maxf, maxf);
    lambda = newi - p.ti;
    V2d P = scale(lambda, tangent);
    va = va - scale(ma, P);
    wa -= Ia * cross(ra, P);
    vb = vb + scale(mb, P);
    wb += Ib * cross(rb, P);
    p.ti = newi;
  }
}

template <typename T>
__device__ __forceinline__ void solve_contact_velocity(VC<T>& vc, V2d& va, double& wa, V2d& vb, double& wb) {
  const V2d normal = dv2(vc.nx, vc.ny), tangent = cross_vs(normal, 1.0);
  const double ma = vc.ma, mb = vc.mb, Ia = vc.Ia, Ib = vc.Ib;
  // tangent constraints first (points in order), then normal
  tangent_point(vc, vc.p[0], tangent, va, wa, vb, wb);
  if (vc.npts == 2) tangent_point(vc, vc.p[1], tangent, va, wa, vb, wb);
  if constexpr (FusedSolve<T>::value) {
    if (vc.npts == 1) {
      VCP<T>& p = vc.p[0];
      const double rax = p.rax, ray = p.ray, rbx = p.rbx, rby = p.rby;
      double dvx, dvy;
      rel_vel_fused(va, wa, vb, wb, rax, ray, rbx, rby, dvx, dvy);
      const double vn = fmad(dvx, normal.x, dvy * normal.y);
      double lambda = -(double)p.nm * (vn - (double)p.bias);
      const double newi = pmax(p.ni + lambda, 0.0);
      lambda = newi - p.ni;
      apply_fused(ma, mb, Ia, Ib, rax, ray, rbx, rby, lambda * normal.x, lambda * normal.y, va, wa, vb, wb);
      p.ni = newi;
    } else if (vc.npts == 2) {  // block solver, total enumeration (contact.py:600-797)
      VCP<T>&cp1 = vc.p[0], &cp2 = vc.p[1];
      const double ax = cp1.ni, ay = cp2.ni;
      const double ra1x = cp1.rax, ra1y = cp1.ray, rb1x = cp1.rbx, rb1y = cp1.rby;
      const double ra2x = cp2.rax, ra2y = cp2.ray, rb2x = cp2.rbx, rb2y = cp2.rby;
      double d1x, d1y, d2x, d2y;
      rel_vel_fused(va, wa, vb, wb, ra1x, ra1y, rb1x, rb1y, d1x, d1y);
      rel_vel_fused(va, wa, vb, wb, ra2x, ra2y, rb2x, rb2y, d2x, d2y);
      const double vn1 = fmad(d1x, normal.x, d1y * normal.y), vn2 = fmad(d2x, normal.x, d2y * normal.y);
      // b = vn - bias - K a, K = [[k11, k12], [k12, k22]]
      const double bx = (vn1 - (double)cp1.bias) - fmad(vc.k11, ax, vc.k12 * ay);
      const double by = (vn2 - (double)cp2.bias) - fmad(vc.k12, ax, vc.k22 * ay);
      const double x1x = -fmad(vc.n11, bx, vc.n12 * by), x1y = -fmad(vc.n21, bx, vc.n22 * by);  // case 1
      const double x2x = -(double)cp1.nm * bx, vn2_2 = fmad(vc.k12, x2x, by);                    // case 2
      const double x3y = -(double)cp2.nm * by, vn1_3 = fmad(vc.k12, x3y, bx);                    // case 3
      const bool c1 = x1x >= 0.0 && x1y >= 0.0;
      const bool c2 = x2x >= 0.0 && vn2_2 >= 0.0;
      const bool c3 = x3y >= 0.0 && vn1_3 >= 0.0;
      const bool c4 = bx >= 0.0 && by >= 0.0;  // case 4: x = 0
      const double xx = c1 ? x1x : (c2 ? x2x : 0.0), xy = c1 ? x1y : (c3 && !c2 ? x3y : 0.0);
      if (c1 || c2 || c3 || c4) {
        const double dx = xx - ax, dy = xy - ay;
        const double p1x = dx * normal.x, p1y = dx * normal.y, p2x = dy * normal.x, p2y = dy * normal.y;
        const double px = p1x + p2x, py = p1y + p2y;
        va.x = fmad(-ma, px, va.x);
        va.y = fmad(-ma, py, va.y);
        wa = fmad(-Ia, fmad(ra1x, p1y, -ra1y * p1x) + fmad(ra2x, p2y, -ra2y * p2x), wa);
        vb.x = fmad(mb, px, vb.x);
        vb.y = fmad(mb, py, vb.y);
        wb = fmad(Ib, fmad(rb1x, p1y, -rb1y * p1x) + fmad(rb2x, p2y, -rb2y * p2x), wb);
        cp1.ni = xx;
        cp2.ni = xy;
      }
    }
  } else if (vc.npts == 1) {
    VCP<T>& p = vc.p[0];
    V2d ra = dv2(p.rax, p.ray), rb = dv2(p.rbx, p.rby);
    V2d dv = vb + cross_sv(wb, rb) - va - cross_sv(wa, ra);
    double vn = dot(dv, normal);
    double lambda = -(double)p.nm * (vn - (double)p.bias);
    double newi = pmax(p.ni + lambda, 0.0);
    lambda = newi - p.ni;
    V2d P = scale(lambda, normal);
    va = va - scale(ma, P);
    wa -= Ia * cross(ra, P);
    vb = vb + scale(mb, P);
    wb += Ib * cross(rb, P);
    p.ni = newi;
  } else if (vc.npts == 2) {  // block solver, total enumeration (contact.py:600-797)
    VCP<T>&cp1 = vc.p[0], &cp2 = vc.p[1];
    V2d a = v2<double>(cp1.ni, cp2.ni);
    V2d ra1 = dv2(cp1.rax, cp1.ray), rb1 = dv2(cp1.rbx, cp1.rby);
    V2d ra2 = dv2(cp2.rax, cp2.ray), rb2 = dv2(cp2.rbx, cp2.rby);
    V2d dv1 = vb + cross_sv(wb, rb1) - va - cross_sv(wa, ra1);
    V2d dv2_ = vb + cross_sv(wb, rb2) - va - cross_sv(wa, ra2);
    double vn1 = dot(dv1, normal), vn2 = dot(dv2_, normal);
    V2d b = v2<double>(vn1 - (double)cp1.bias, vn2 - (double)cp2.bias);
    M22<double> K;
    K.c1x = vc.k11; K.c1y = vc.k12; K.c2x = vc.k12; K.c2y = vc.k22;
    M22<double> N;
    N.c1x = vc.n11; N.c1y = vc.n21; N.c2x = vc.n12; N.c2y = vc.n22;
    b = b - mmul(K, a);
    // The four cases are evaluated side by side and the first valid one is applied once:
    // the same values as trying them in turn, without a divergent path per case.
    const V2d x1 = -mmul(N, b);                                    // case 1
    const V2d x2 = v2<double>(-(double)cp1.nm * b.x, 0.0);         // case 2
    const double vn2_2 = K.c1y * x2.x + b.y;
    const V2d x3 = v2<double>(0.0, -(double)cp2.nm * b.y);         // case 3
    const double vn1_3 = K.c2x * x3.y + b.x;
    const bool c1 = x1.x >= 0.0 && x1.y >= 0.0;
    const bool c2 = x2.x >= 0.0 && vn2_2 >= 0.0;
    const bool c3 = x3.y >= 0.0 && vn1_3 >= 0.0;
    const bool c4 = b.x >= 0.0 && b.y >= 0.0;                       // case 4: x = 0
    const V2d x = c1 ? x1 : (c2 ? x2 : (c3 ? x3 : v2<double>(0.0, 0.0)));
    if (c1 || c2 || c3 || c4) apply_block(vc, x, a, normal, va, wa, vb, wb);
  }
}


// ---- software-pipelined Gauss-Seidel sweeps ----
// Each lane walks its island's contacts in order.  While contact k is being
// solved, the record of contact k+1 and the velocities (positions) of its
// bodies that contact k does not touch are already in flight; bodies shared
// with contact k are forwarded from registers.  This keeps the exact
// sequential order and hides most of the memory latency of the chain.
template <typename T>
__device__ __forceinline__ void vel_of(const Dev<T>& d, int bstart, int i, V2d& v, double& w) {
  if (i >= 0) {
    const PV<T>* p = &d.pv[bstart + i];
    v = v2<double>(p->vx, p->vy);
    w = p->w;
  } else {  // static bodies have zero velocity (checked at upload)
    v = v2<double>(0.0, 0.0);
    w = 0.0;
  }
}

template <typename T>
__device__ __forceinline__ void warm_point(const VC<T>& vc, const VCP<T>& p, V2d normal, V2d tangent, V2d& va,
                                           double& wa, V2d& vb, double& wb) {
  if constexpr (FusedSolve<T>::value) {
    const double Px = fmad(p.ni, normal.x, p.ti * tangent.x), Py = fmad(p.ni, normal.y, p.ti * tangent.y);
    apply_fused(vc.ma, vc.mb, vc.Ia, vc.Ib, p.rax, p.ray, p.rbx, p.rby, Px, Py, va, wa, vb, wb);
    return;
  }
  V2d P = scale(p.ni, normal) + scale(p.ti, tangent);
  wa -= (double)vc.Ia * cross(dv2(p.rax, p.ray), P);
  va = va - scale((double)vc.ma, P);
  wb += (double)vc.Ib * cross(dv2(p.rbx, p.rby), P);
  vb = vb + scale((double)vc.mb, P);
}

template <typename T>
__device__ __forceinline__ void warm_contact(const VC<T>& vc, V2d& va, double& wa, V2d& vb, double& wb) {
  V2d normal = dv2(vc.nx, vc.ny), tangent = cross_vs(normal, 1.0);
  if (vc.npts >= 1) warm_point(vc, vc.p[0], normal, tangent, va, wa, vb, wb);
  if (vc.npts == 2) warm_point(vc, vc.p[1], normal, tangent, va, wa, vb, wb);
}

template <typename T, bool kWarm>
__device__ void velocity_sweep(const Dev<T>& d, int bstart, int cstart, int nci) {
  if (nci <= 0) return;
  VC<T> A = d.vc[cstart];
  V2d va, vb;
  double wa, wb;
  vel_of(d, bstart, A.ia, va, wa);
  vel_of(d, bstart, A.ib, vb, wb);
  int2 nidx = nci > 1 ? d.i_cidx[cstart + 1] : make_int2(0, 0);
  for (int k = 0; k < nci; ++k) {
    const bool more = k + 1 < nci;
    VC<T> B;
    int2 idx2 = make_int2(0, 0);
    V2d pva = v2<double>(0, 0), pvb = v2<double>(0, 0);
    double pwa = 0.0, pwb = 0.0;
    if (more) {
      B = d.vc[cstart + k + 1];
      if (k + 2 < nci) idx2 = d.i_cidx[cstart + k + 2];
      if (nidx.x != A.ia && nidx.x != A.ib) vel_of(d, bstart, nidx.x, pva, pwa);
      if (nidx.y != A.ia && nidx.y != A.ib) vel_of(d, bstart, nidx.y, pvb, pwb);
    }
    if (kWarm) warm_contact(A, va, wa, vb, wb);
    else solve_contact_velocity(A, va, wa, vb, wb);
    store_vel(d, bstart, A.ia, va, wa);
    store_vel(d, bstart, A.ib, vb, wb);
    if (!kWarm) {
      VCP<T>* dst = d.vc[cstart + k].p;
      dst[0].ni = A.p[0].ni;
      dst[0].ti = A.p[0].ti;
      if (A.npts == 2) {
        dst[1].ni = A.p[1].ni;
        dst[1].ti = A.p[1].ti;
      }
    }
    if (more) {
      V2d nva, nvb;
      double nwa, nwb;
      if (nidx.x == A.ia) { nva = va; nwa = wa; }
      else if (nidx.x == A.ib) { nva = vb; nwa = wb; }
      else { nva = pva; nwa = pwa; }
      if (nidx.y == A.ia) { nvb = va; nwb = wa; }
      else if (nidx.y == A.ib) { nvb = vb; nwb = wb; }
      else { nvb = pvb; nwb = pwb; }
      A = B;
      va = nva; wa = nwa; vb = nvb; wb = nwb;
      nidx = idx2;
    }
  }
}

template <typename T>
__device__ __forceinline__ void pos_of(const Dev<T>& d, int bstart, int i, const PC<T>& pc, double& x, double& y,
                                       double& a) {
  if (i >= 0) {
    const PV<T>* p = &d.pv[bstart + i];
    x = p->cx;
    y = p->cy;
    a = p->a;
  } else {
    x = pc.scx;
    y = pc.scy;
    a = pc.sa;
  }
}

// One contact of solve_position_constraints (contact.py:809-864).
template <typename T>
__device__ __forceinline__ void solve_point_position(const PC<T>& pc, int j, double& cax, double& cay, double& aa,
                                                     double& cbx, double& cby, double& ab, T& min_sep) {
  {
    Rot<T> qa = rot_of<T>(aa), qb = rot_of<T>(ab);
    V2<T> dcb = v2<T>((T)(cbx - cax), (T)(cby - cay));
    XF<T> xfa, xfb;
    xfa.q = qa;
    xfa.p = -rmul(qa, v2<T>(pc.lcax, pc.lcay));
    xfb.q = qb;
    xfb.p = dcb - rmul(qb, v2<T>(pc.lcbx, pc.lcby));
    // PositionSolverManifold (contact_util.py): circles, face A and face B share one sequence of
    // operations with the roles of the two transforms as selects (the same arithmetic per type
    // as three branches, without a divergent path per type).  The reference point lpp goes
    // through X, the clip point through Y: circles and face A: X = A, Y = B; face B swapped.
    const bool circ = pc.type == M_CIRCLES, face_b = !circ && pc.type != M_FACE_A;
    const V2<T> lpj = j == 0 ? v2<T>(pc.lp0x, pc.lp0y) : v2<T>(pc.lp1x, pc.lp1y);
    const V2<T> ln = v2<T>(pc.lnx, pc.lny), lpp = v2<T>(pc.lpx, pc.lpy);
    XF<T> X, Y;
    X.q.c = face_b ? xfb.q.c : xfa.q.c; X.q.s = face_b ? xfb.q.s : xfa.q.s;
    X.p.x = face_b ? xfb.p.x : xfa.p.x; X.p.y = face_b ? xfb.p.y : xfa.p.y;
    Y.q.c = face_b ? xfa.q.c : xfb.q.c; Y.q.s = face_b ? xfa.q.s : xfb.q.s;
    Y.p.x = face_b ? xfa.p.x : xfb.p.x; Y.p.y = face_b ? xfa.p.y : xfb.p.y;
    const V2<T> plane = xmul(X, lpp), clip = xmul(Y, lpj);  // (a circle manifold has one point)
    const V2<T> dd = clip - plane;
    const V2<T> nrm = circ ? normalized(dd) : rmul(X.q, ln);
    const T sep = dot(dd, nrm) - pc.ra - pc.rb;
    const V2<T> point = circ ? scale((T)0.5, plane + clip) : clip;
    const V2<T> normal = face_b ? -nrm : nrm;
    V2<T> ra = point, rb = point - dcb;
    min_sep = pmin(min_sep, sep);
    T C = pclamp((T)BAUMGARTE * (sep + (T)LINEAR_SLOP), -(T)MAX_LINEAR_CORRECTION, (T)0);
    T rna = sq(cross(ra, normal)), rnb = sq(cross(rb, normal));
    T K = pc.ma + pc.mb + (pc.Ia * rna) + (pc.Ib * rnb);
    T impulse = K > (T)0 ? -C / K : (T)0;
    V2<T> P = scale(impulse, normal);
    cax -= (double)(pc.ma * P.x);
    cay -= (double)(pc.ma * P.y);
    aa -= (double)(pc.Ia * cross(ra, P));
    cbx += (double)(pc.mb * P.x);
    cby += (double)(pc.mb * P.y);
    ab += (double)(pc.Ib * cross(rb, P));
  }
}
// The points are solved in order, each from the positions the previous one left
// (written out rather than as a loop: a loop here makes the compiler copy the
// team path's record at every branch of its pass).
template <typename T>
__device__ __forceinline__ void solve_contact_position(const PC<T>& pc, double& cax, double& cay, double& aa,
                                                       double& cbx, double& cby, double& ab, T& min_sep) {
  if (pc.npts > 0) solve_point_position(pc, 0, cax, cay, aa, cbx, cby, ab, min_sep);
  if (pc.npts > 1) solve_point_position(pc, 1, cax, cay, aa, cbx, cby, ab, min_sep);
}

template <typename T>
__device__ bool position_sweep(const Dev<T>& d, int bstart, int cstart, int nci) {
  T min_sep = 0;
  if (nci > 0) {
    PC<T> A = d.pc[cstart];
    double ax, ay, aa, bx, by, ab;
    pos_of(d, bstart, A.ia, A, ax, ay, aa);
    pos_of(d, bstart, A.ib, A, bx, by, ab);
    int2 nidx = nci > 1 ? d.i_cidx[cstart + 1] : make_int2(0, 0);
    for (int k = 0; k < nci; ++k) {
      const bool more = k + 1 < nci;
      PC<T> B;
      int2 idx2 = make_int2(0, 0);
      double p0x = 0, p0y = 0, p0a = 0, p1x = 0, p1y = 0, p1a = 0;
      if (more) {
        B = d.pc[cstart + k + 1];
        if (k + 2 < nci) idx2 = d.i_cidx[cstart + k + 2];
        if (nidx.x >= 0 && nidx.x != A.ia && nidx.x != A.ib) {
          const PV<T>* p = &d.pv[bstart + nidx.x];
          p0x = p->cx; p0y = p->cy; p0a = p->a;
        }
        if (nidx.y >= 0 && nidx.y != A.ia && nidx.y != A.ib) {
          const PV<T>* p = &d.pv[bstart + nidx.y];
          p1x = p->cx; p1y = p->cy; p1a = p->a;
        }
      }
      solve_contact_position(A, ax, ay, aa, bx, by, ab, min_sep);
      store_pos(d, bstart, A.ia, ax, ay, aa);
      store_pos(d, bstart, A.ib, bx, by, ab);
      if (more) {
        double nax, nay, naa, nbx, nby, nab;
        if (nidx.x < 0) { nax = B.scx; nay = B.scy; naa = B.sa; }
        else if (nidx.x == A.ia) { nax = ax; nay = ay; naa = aa; }
        else if (nidx.x == A.ib) { nax = bx; nay = by; naa = ab; }
        else { nax = p0x; nay = p0y; naa = p0a; }
        if (nidx.y < 0) { nbx = B.scx; nby = B.scy; nab = B.sa; }
        else if (nidx.y == A.ia) { nbx = ax; nby = ay; nab = aa; }
        else if (nidx.y == A.ib) { nbx = bx; nby = by; nab = ab; }
        else { nbx = p1x; nby = p1y; nab = p1a; }
        A = B;
        ax = nax; ay = nay; aa = naa; bx = nbx; by = nby; ab = nab;
        nidx = idx2;
      }
    }
  }
  return min_sep >= -(T)(3.0 * LINEAR_SLOP);
}

// Per-island constants.
struct IslandK {
  int w, b0, f0, cbase, wflags, pvb;
  double dt, dt_ratio;
  V2d grav;
  bool warm;
};

template <typename T>
__device__ __forceinline__ IslandK island_k(const Dev<T>& d, int w, int bstart) {
  IslandK c;
  c.w = w;
  c.b0 = d.w_body_off[w];
  c.f0 = d.w_fix_off[w];
  c.cbase = d.w_con_base[w];
  c.wflags = d.w_flags[w];
  c.pvb = bstart - c.b0;  // island body i is world-segment entry pvb + i
  c.dt = d.dt;
  c.dt_ratio = d.w_inv_dt0[w] * d.dt;
  c.grav = v2<double>(d.w_grav[2 * w], d.w_grav[2 * w + 1]);
  c.warm = (c.wflags & WF_WARM) != 0;
  return c;
}

// 1. integrate velocities of island body i (island.py:71-95)
template <typename T>
__device__ __forceinline__ PV<T> integrate_velocity(const Dev<T>& d, const IslandK& c, int bstart, int i) {
  const int gb = c.b0 + d.i_bodies[bstart + i];
  double2 cc = d.b_c[gb];
  double a = d.b_a[gb];
  d.b_c0[gb] = cc;
  d.b_a0[gb] = a;
  V2d v = d.b_v[gb];
  double wv = d.b_w[gb];
  if (d.b_type[gb] == BT_DYNAMIC) {
    const double dt = c.dt;
    V2<T> f = d.b_f[gb];
    V2d acc = scale((double)d.b_gs[gb], c.grav) + scale((double)d.b_im[gb], dv2(f.x, f.y));
    v = v + scale(dt, acc);
    wv += dt * (double)d.b_iI[gb] * (double)d.b_t[gb];
    v = scale(pclamp(1.0 - dt * (double)d.b_ld[gb], 0.0, 1.0), v);
    wv *= pclamp(1.0 - dt * (double)d.b_ad[gb], 0.0, 1.0);
  }
  PV<T> p;
  p.cx = cc.x;
  p.cy = cc.y;
  p.a = a;
  p.vx = v.x;
  p.vy = v.y;
  p.w = wv;
  return p;
}

// 2. ContactSolver.__init__ + initialize_velocity_constraints for island
// contact k (contact.py:376-496).  Reads the island's integrated state from
// d.pv; writes the constraint records and turns i_cidx island-local.
template <typename T>
__device__ void init_contact(const Dev<T>& d, const IslandK& c, int bstart, int cstart, int k, int& ia_out,
                             int& ib_out, int rka = 0, int rkb = 0) {
  const int b0 = c.b0, f0 = c.f0, pvb = c.pvb;
  const int gs_ = c.cbase + d.i_contacts[cstart + k];
  int2 cix = d.i_cidx[cstart + k];
  const uint2 fbx = d.i_fb[cstart + k];
  const int ia = cix.x >= 0 ? cix.x - pvb : cix.x, ib = cix.y >= 0 ? cix.y - pvb : cix.y;
  ia_out = ia;
  ib_out = ib;
  const int fa = f0 + (int)(fbx.x & 0xFFFFu), fb = f0 + (int)(fbx.x >> 16);
  const int ba = b0 + (int)(fbx.y & 0xFFFFu), bb = b0 + (int)(fbx.y >> 16);
  const int st = d.c_state[gs_];
  const int mtype = (st >> CS_MTYPE_SHIFT) & 3, pc = (st >> CS_PC_SHIFT) & 3;
  V4<T> lnlp = d.c_lnlp[gs_], pts = d.c_pts[gs_];
  const V4<double> imp = d.c_imp[gs_];
  V2<T> mat = d.c_mat[gs_];
  VC<T> vc;
  PC<T> pcn;
  vc.friction = mat.x;
  vc.restitution = mat.y;
  vc.ia = ia;
  vc.ib = ib;
  vc.ma = d.b_im[ba];
  vc.mb = d.b_im[bb];
  vc.Ia = d.b_iI[ba];
  vc.Ib = d.b_iI[bb];
  vc.npts = pc;
  vc.nxt = 0u;
  vc.rka = rka;
  vc.rkb = rkb;
  const bool warm = c.warm;
  const double dt_ratio = c.dt_ratio;
  vc.p[0].ni = warm ? dt_ratio * (double)imp.x : (double)imp.x;
  vc.p[0].ti = warm ? dt_ratio * (double)imp.y : (double)imp.y;
  vc.p[1].ni = warm ? dt_ratio * (double)imp.z : (double)imp.z;
  vc.p[1].ti = warm ? dt_ratio * (double)imp.w : (double)imp.w;
  double2 lca = d.b_lc[ba], lcb = d.b_lc[bb];
  pcn.lp0x = pts.x; pcn.lp0y = pts.y; pcn.lp1x = pts.z; pcn.lp1y = pts.w;
  pcn.lnx = lnlp.x; pcn.lny = lnlp.y; pcn.lpx = lnlp.z; pcn.lpy = lnlp.w;
  pcn.ra = (T)d.f_rad64[fa];
  pcn.rb = (T)d.f_rad64[fb];
  pcn.ma = vc.ma; pcn.mb = vc.mb; pcn.Ia = vc.Ia; pcn.Ib = vc.Ib;
  pcn.lcax = (T)lca.x; pcn.lcay = (T)lca.y; pcn.lcbx = (T)lcb.x; pcn.lcby = (T)lcb.y;
  pcn.type = mtype;
  pcn.npts = pc;
  pcn.ia = ia;
  pcn.ib = ib;
  pcn.rka = rka;
  pcn.rkb = rkb;
  pcn.nxt = pcn.pad = 0u;
  double cax, cay, aa, cbx, cby, ab;
  load_pos(d, bstart, b0, ia, cax, cay, aa);
  load_pos(d, bstart, b0, ib, cbx, cby, ab);
  pcn.scx = ia < 0 ? cax : cbx;
  pcn.scy = ia < 0 ? cay : cby;
  pcn.sa = ia < 0 ? aa : ab;
  d.i_cidx[cstart + k] = make_int2(ia, ib);  // island-local from here on
  V2d va, vb;
  double wa, wb;
  load_vel(d, bstart, b0, ia, va, wa);
  load_vel(d, bstart, b0, ib, vb, wb);
  // The constraint geometry is computed in FP64 in both pipelines: it runs once
  // per contact and step, and FP32 arithmetic here (rotations, world manifold,
  // effective masses) would set the FP32 pipeline's one-step error.  Normal and
  // lever arms are stored in FP64 (VCP), effective masses and bias as T.
  const Rot<double> qa = rot_of<double>(aa), qb = rot_of<double>(ab);
  const V2d dcb = v2<double>(cbx - cax, cby - cay);
  XF<double> xfa, xfb;
  xfa.q = qa;
  xfa.p = -rmul(qa, v2<double>(lca.x, lca.y));
  xfb.q = qb;
  xfb.p = dcb - rmul(qb, v2<double>(lcb.x, lcb.y));
  V2d w0 = v2<double>(0, 0), w1 = v2<double>(0, 0);
  const V2d normal = world_manifold<double>(mtype, pc, dv2(lnlp.x, lnlp.y), dv2(lnlp.z, lnlp.w), dv2(pts.x, pts.y),
                                            dv2(pts.z, pts.w), xfa, d.f_rad64[fa], xfb, d.f_rad64[fb], w0, w1);
  vc.nx = normal.x;
  vc.ny = normal.y;
  const T ma = vc.ma, mb = vc.mb, Ia = vc.Ia, Ib = vc.Ib;
  const V2d nd = dv2(vc.nx, vc.ny);
  auto init_point = [&](VCP<T>& p, V2d wp) {
    const double mad = ma, mbd = mb, Iad = Ia, Ibd = Ib;
    V2d ra = wp, rb = wp - dcb;
    double rna = sq(cross(ra, normal)), rnb = sq(cross(rb, normal));
    double kn = mad + mbd + Iad * rna + Ibd * rnb;
    V2d tangent = cross_vs(normal, 1.0);
    double rta = sq(cross(ra, tangent)), rtb = sq(cross(rb, tangent));
    double kt = mad + mbd + Iad * rta + Ibd * rtb;
    p.rax = ra.x; p.ray = ra.y; p.rbx = rb.x; p.rby = rb.y;
    double vrel = dot(nd, vb + cross_sv(wb, dv2(p.rbx, p.rby)) - va - cross_sv(wa, dv2(p.rax, p.ray)));
    p.nm = kn > 0.0 ? (T)(1.0 / kn) : (T)0;
    p.tm = kt > 0.0 ? (T)(1.0 / kt) : (T)0;
    p.bias = vrel < -VELOCITY_THRESHOLD ? (T)(-(double)vc.restitution * vrel) : (T)0;
    p.pad = 0;
  };
  vc.p[1].rax = vc.p[1].ray = vc.p[1].rbx = vc.p[1].rby = 0.0;
  vc.p[1].nm = vc.p[1].tm = vc.p[1].bias = vc.p[1].pad = (T)0;
  if (pc >= 1) init_point(vc.p[0], w0);
  if (pc >= 2) init_point(vc.p[1], w1);
  vc.k11 = 1; vc.k12 = 0; vc.k22 = 1; vc.n11 = 1; vc.n12 = 0; vc.n21 = 0; vc.n22 = 1;
  if (pc == 2) {
    V2d ra1 = dv2(vc.p[0].rax, vc.p[0].ray), rb1 = dv2(vc.p[0].rbx, vc.p[0].rby);
    V2d ra2 = dv2(vc.p[1].rax, vc.p[1].ray), rb2 = dv2(vc.p[1].rbx, vc.p[1].rby);
    const double mad = ma, mbd = mb, Iad = Ia, Ibd = Ib;
    double rn1a = cross(ra1, nd), rn1b = cross(rb1, nd);
    double rn2a = cross(ra2, nd), rn2b = cross(rb2, nd);
    double k11 = mad + mbd + Iad * rn1a * rn1a + Ibd * rn1b * rn1b;
    double k22 = mad + mbd + Iad * rn2a * rn2a + Ibd * rn2b * rn2b;
    double k12 = mad + mbd + Iad * rn1a * rn2a + Ibd * rn1b * rn2b;
    if (sq(k11) < 1000.0 * (k11 * k22 - sq(k12))) {
      vc.k11 = k11; vc.k12 = k12; vc.k22 = k22;
      double det = k11 * k22 - k12 * k12;
      if (det != 0.0) det = 1.0 / det;
      vc.n11 = det * k22; vc.n21 = -det * k12; vc.n12 = -det * k12; vc.n22 = det * k11;
    } else {
      vc.npts = 1;  // quirk Q6: keep only the first point
    }
  }
  d.vc[cstart + k] = vc;
  d.pc[cstart + k] = pcn;
}

// 5. store_impulses (contact.py:801-806): zip(manifold points, constraint points)
template <typename T>
__device__ __forceinline__ void store_impulse(const Dev<T>& d, const IslandK& c, int cstart, int k) {
  const VC<T>& vc = d.vc[cstart + k];
  const int gs_ = c.cbase + d.i_contacts[cstart + k];
  const int pc = (d.c_state[gs_] >> CS_PC_SHIFT) & 3;
  const int n = pc < vc.npts ? pc : vc.npts;
  V4<double> imp = d.c_imp[gs_];
  if (n >= 1) { imp.x = vc.p[0].ni; imp.y = vc.p[0].ti; }
  if (n >= 2) { imp.z = vc.p[1].ni; imp.w = vc.p[1].ti; }
  d.c_imp[gs_] = imp;
  if (d.ev_mask & (1 << EV_POST_SOLVE)) {  // Island.report (island.py:245-252): the points' impulses
    const int e = ev_open(d, EV_POST_SOLVE);
    if (e >= 0) {
      ev_write(d, e, c.w, EV_POST_SOLVE, 2, cstart + k - c.cbase, gs_, d.c_state[gs_]);
      double* ed = d.ev_d + (size_t)e * B2G_EV_ND;
      int* ei = d.ev_i + (size_t)e * B2G_EV_NI;
      // the impulses handed to the listener: one per velocity-constraint point (quirk Q6 may
      // have dropped the second), in the record's old-manifold fields
      ei[B2G_EI_OLD_PC] = vc.npts;
      ed[B2G_ED_OLD + B2G_ED_P0_NI] = vc.p[0].ni;
      ed[B2G_ED_OLD + B2G_ED_P0_TI] = vc.p[0].ti;
      ed[B2G_ED_OLD + B2G_ED_P1_NI] = vc.npts >= 2 ? vc.p[1].ni : 0.0;
      ed[B2G_ED_OLD + B2G_ED_P1_TI] = vc.npts >= 2 ? vc.p[1].ti : 0.0;
    }
  }
}

// 6. integrate positions (island.py:120-138)
template <typename T>
__device__ __forceinline__ void integrate_position(PV<T>& p, double dt) {
  V2d v = v2<double>(p.vx, p.vy);
  double wv = p.w;
  V2d tr = scale(dt, v);
  if (dot(tr, tr) > MAX_TRANSLATION * MAX_TRANSLATION) {
    double ratio = MAX_TRANSLATION / sqrt(sq(tr.x) + sq(tr.y));
    v = scale(ratio, v);
  }
  double rotation = dt * wv;
  if (sq(rotation) > MAX_ROTATION * MAX_ROTATION) {
    double ratio = MAX_ROTATION / fabs(rotation);
    wv *= ratio;
  }
  V2d dc = scale(dt, v);
  p.cx = p.cx + dc.x;
  p.cy = p.cy + dc.y;
  p.a += dt * wv;
  p.vx = v.x;
  p.vy = v.y;
  p.w = wv;
}

// 8. copy back and Body._synchronize_transform (island.py:150-159, body.py:906-908)
template <typename T>
__device__ __forceinline__ void write_back(const Dev<T>& d, const IslandK& c, int bstart, int i, const PV<T>& p) {
  const int gb = c.b0 + d.i_bodies[bstart + i];
  d.b_c[gb] = make_double2(p.cx, p.cy);
  d.b_a[gb] = p.a;
  d.b_v[gb] = v2<double>(p.vx, p.vy);
  d.b_w[gb] = p.w;
  double sn, cs;
  sincos(p.a, &sn, &cs);
  double2 lc = d.b_lc[gb];
  d.b_xf[gb] = make_double4(p.cx - (cs * lc.x + (-sn) * lc.y), p.cy - (sn * lc.x + cs * lc.y), cs, sn);
  // non-finite state is reported (B2G_E_NAN at the next sync); the step's kernels still end
  // normally, since every loop on this path is bounded by counts, not by values
  if (!(isfinite(p.cx) && isfinite(p.cy) && isfinite(p.a) && isfinite(p.vx) && isfinite(p.vy) && isfinite(p.w)))
    atomicOr(&d.w_err[c.w], WERR_NAN);
}

// 9. sleep timer of island body i (island.py:161-179, quirk Q7: the stale
// loop variables); returns the body's contribution to min_sleep_time.
template <typename T>
__device__ __forceinline__ double sleep_body(const Dev<T>& d, const IslandK& c, int bstart, int i, int last_allow) {
  const int gb = c.b0 + d.i_bodies[bstart + i];
  V2d v = d.b_v[gb];
  double wv = d.b_w[gb];
  if (!last_allow || sq(wv) > ANGULAR_SLEEP_TOL * ANGULAR_SLEEP_TOL ||
      sq(v.x) + sq(v.y) > LINEAR_SLEEP_TOL * LINEAR_SLEEP_TOL) {
    d.b_sleep[gb] = 0.0;
    return 0.0;
  }
  double t = d.b_sleep[gb] + c.dt;
  d.b_sleep[gb] = t;
  return t;
}

template <typename T>
__device__ __forceinline__ void sleep_last(const Dev<T>& d, const IslandK& c, int bstart, int nbi) {
  const int gb = c.b0 + d.i_bodies[bstart + nbi - 1];
  d.b_sleep[gb] = 0.0;
  d.b_v[gb] = v2<double>(0, 0);
  d.b_w[gb] = 0.0;
  d.b_f[gb] = v2<T>(0, 0);
  d.b_t[gb] = (T)0;
  d.b_flags[gb] &= ~BF_AWAKE;
}

// Island timing marks (b2g_island_times): when the velocity iterations ended, and how many
// position iterations ran before the early exit.
struct SolveMarks {
  long long t_vel;
  int npos;
};

// The island's positions and velocities as the joint solver reads and writes them: island
// bodies in pv, bodies outside the island's body list (static ones) read from the world and
// never written.
template <typename T>
struct IslandIO {
  const Dev<T>* d;
  int pvb, b0;
  __device__ __forceinline__ void pos(int i, V2d& c, double& a) const {
    double cx, cy;
    load_pos(*d, pvb, b0, i, cx, cy, a);
    c = v2<double>(cx, cy);
  }
  __device__ __forceinline__ void vel(int i, V2d& v, double& w) const { load_vel(*d, pvb, b0, i, v, w); }
  __device__ __forceinline__ void set_pos(int i, V2d c, double a) const { store_pos(*d, pvb, i, c.x, c.y, a); }
  __device__ __forceinline__ void set_vel(int i, V2d v, double w) const { store_vel(*d, pvb, i, v, w); }
};

// The island's joint k (an entry of i_joints): its record and its bodies.
template <typename T>
__device__ __forceinline__ JRec island_joint(const Dev<T>& d, int b0, int j0, int e, JBodies& B) {
  const int jl = d.i_joints[e];
  JRec J;
  J.d = d.j_d + (size_t)(j0 + jl) * B2G_JOINT_ND;
  J.i = d.j_i + (size_t)(j0 + jl) * B2G_JOINT_NI;
  const int4 ix = d.i_jidx[e];
  B.idx[0] = ix.x;
  B.idx[1] = ix.y;
  B.idx[2] = ix.z;
  B.idx[3] = ix.w;
  const bool gear = J.i[B2G_JI_TYPE] == B2G_JT_GEAR;  // only gear joints have bodies c, d
  const int bb = J.i[B2G_JI_BODY_B];
  const int slot[4] = {J.i[B2G_JI_BODY_A] < 0 ? bb : J.i[B2G_JI_BODY_A], bb, gear ? J.i[B2G_JI_BODY_C] : bb,
                       gear ? J.i[B2G_JI_BODY_D] : bb};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const double4 m = d.j_mass[4 * (size_t)(j0 + jl) + q];
    B.im[q] = m.x;
    B.iI[q] = m.y;
    B.mass[q] = m.z;
    const double2 lc = d.b_lc[b0 + slot[q]];
    B.lc[q] = v2<double>(lc.x, lc.y);
  }
  return J;
}

// Island.solve on one lane (island.py:65-179) for an island with joints, the joints in their
// places (island.py:107-114 and 141-152).  Runs in k_solve_joints, so that the joint solver's
// registers stay out of k_solve.
template <typename T>
__device__ void solve_island_joints(const Dev<T>& d, int w, int bstart, int nbi, int cstart, int nci,
                                    int last_allow, SolveMarks& mk) {
  const IslandK c = island_k(d, w, bstart);
  const int2 jr = d.i_jr[bstart];
  const int je = jr.x, nje = jr.y, j0 = d.w_joint_off[w];
  const IslandIO<T> io{&d, bstart, c.b0};
  JStep jst;
  jst.dt = c.dt;
  jst.inv_dt = d.inv_dt;
  jst.dt_ratio = c.dt_ratio;
  jst.warm = c.warm;
  // 1 + 2. velocities integrated and constraint records built by k_prep_bodies / k_prep_contacts
  // 3. warm_start (contact.py:498-525), the joints' setup, and 4. velocity iterations
  // (contact.py:527-799), joints before contacts
  if (c.warm) velocity_sweep<T, true>(d, bstart, cstart, nci);
  for (int k = 0; k < nje; ++k) {
    JBodies B;
    const JRec J = island_joint(d, c.b0, j0, je + k, B);
    if (!joint_init(J, B, io, jst)) atomicOr(&d.w_err[w], WERR_JOINT);
  }
  for (int it = 0; it < d.vel_iters; ++it) {
    for (int k = 0; k < nje; ++k) {
      JBodies B;
      const JRec J = island_joint(d, c.b0, j0, je + k, B);
      joint_solve_velocity(J, B, io, jst);
    }
    velocity_sweep<T, false>(d, bstart, cstart, nci);
  }
  if (d.i_time) mk.t_vel = (long long)wall_clock64();
  for (int k = 0; k < nci; ++k) store_impulse(d, c, cstart, k);
  for (int i = 0; i < nbi; ++i) {
    PV<T> p = d.pv[bstart + i];
    integrate_position(p, c.dt);
    d.pv[bstart + i] = p;
  }
  // 7. position iterations with early exit (contact.py:809-864), contacts then every joint
  bool position_solved = false;
  for (int it = 0; it < d.pos_iters; ++it) {
    if (d.i_time) mk.npos = it + 1;
    const bool contacts_okay = position_sweep(d, bstart, cstart, nci);
    bool joints_okay = true;
    for (int k = 0; k < nje; ++k) {
      JBodies B;
      const JRec J = island_joint(d, c.b0, j0, je + k, B);
      const bool ok = joint_solve_position(J, B, io);
      joints_okay = joints_okay && ok;
    }
    if (contacts_okay && joints_okay) {
      position_solved = true;
      break;
    }
  }
  for (int i = 0; i < nbi; ++i) write_back(d, c, bstart, i, d.pv[bstart + i]);
  if (c.wflags & WF_ALLOW_SLEEP) {
    double min_sleep = 1.0e300;
    for (int i = 0; i < nbi; ++i) min_sleep = pmin(min_sleep, sleep_body(d, c, bstart, i, last_allow));
    if (min_sleep >= TIME_TO_SLEEP && position_solved && nbi > 0) sleep_last(d, c, bstart, nbi);
  }
}

// Island.solve on one lane (island.py:65-179).
template <typename T>
__device__ void solve_island(const Dev<T>& d, int w, int bstart, int nbi, int cstart, int nci, int last_allow,
                             SolveMarks& mk) {
  const IslandK c = island_k(d, w, bstart);
  // 1 + 2. velocities integrated and constraint records built by k_prep_bodies / k_prep_contacts
  // 3. warm_start (contact.py:498-525) and 4. velocity iterations (contact.py:527-799)
  if (c.warm) velocity_sweep<T, true>(d, bstart, cstart, nci);
  for (int it = 0; it < d.vel_iters; ++it) velocity_sweep<T, false>(d, bstart, cstart, nci);
  if (d.i_time) mk.t_vel = (long long)wall_clock64();
  for (int k = 0; k < nci; ++k) store_impulse(d, c, cstart, k);
  for (int i = 0; i < nbi; ++i) {
    PV<T> p = d.pv[bstart + i];
    integrate_position(p, c.dt);
    d.pv[bstart + i] = p;
  }
  // 7. position iterations with early exit (contact.py:809-864)
  bool position_solved = false;
  for (int it = 0; it < d.pos_iters; ++it) {
    if (d.i_time) mk.npos = it + 1;
    if (position_sweep(d, bstart, cstart, nci)) {
      position_solved = true;
      break;
    }
  }
  for (int i = 0; i < nbi; ++i) write_back(d, c, bstart, i, d.pv[bstart + i]);
  if (c.wflags & WF_ALLOW_SLEEP) {
    double min_sleep = 1.0e300;
    for (int i = 0; i < nbi; ++i) min_sleep = pmin(min_sleep, sleep_body(d, c, bstart, i, last_allow));
    if (min_sleep >= TIME_TO_SLEEP && position_solved && nbi > 0) sleep_last(d, c, bstart, nbi);
  }
}

// Does the island starting at island-body bstart (global) have joints?
template <typename T>
__device__ __forceinline__ bool island_has_joints(const Dev<T>& d, int w, int bstart) {
  return d.w_joint_off && d.w_joint_off[w + 1] > d.w_joint_off[w] && d.i_jr[bstart].y > 0;
}

// ---------------------------------------------------------------------------
// Island.solve by a team of lanes (the path for large islands)
//
// A team is 8, 16, 32 or 64 lanes of one wavefront (by the island's size
// class); a wavefront carries 64 / T islands of one class.  The island search
// orders each island's contacts by dependency level within one sweep; lane t
// of a team owns entries t, t + T, t + 2T, ... of that order.  A contact's
// solve at sweep s must see every earlier solve (in the reference's
// Gauss-Seidel order) that touches one of its dynamic bodies, and nothing
// later.  Each dynamic body X keeps a counter of completed solves that touched
// it; the contact is ready when, for each of its dynamic bodies, the counter
// equals its rank in that body's sequence (s * deg(X) + the number of the
// body's contacts before it in the island list, from the island search).
// Ready lanes run in the same pass, which is safe because at most one pending
// solve per body can be ready.  A lane takes its solves in order of
// s * period + level (sweep_period): every dependency goes to a larger key, so
// the unfinished solve with the smallest key has only finished predecessors
// and is at the head of its lane, and every pass makes progress; passes are
// still capped.  Body velocities, then positions, live in LDS; the constraint
// records stay in HBM, one per lane in registers.  The arithmetic per contact
// is the same code as the one-lane path, so results are identical.
// ---------------------------------------------------------------------------

// The lanes of one wavefront exchange body state through LDS between passes.
// LDS instructions of a wavefront execute in issue order, so wavefront-scope
// ordering (which emits no waits) is all a pass boundary needs.
__device__ __forceinline__ int lds_ld(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}
__device__ __forceinline__ void lds_st(int* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}
__device__ __forceinline__ void pass_fence() {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
}
// Phase boundary inside a one-wavefront block: makes the team's LDS and global
// writes visible to the other lanes (no s_barrier, so divergent teams are fine).
__device__ __forceinline__ void team_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  __builtin_amdgcn_wave_barrier();
}
__device__ __forceinline__ bool wave_any(bool p) {
  return __ballot(p) != 0;
}
template <typename X>
__device__ __forceinline__ X team_min(X v, int T) {
  for (int off = T >> 1; off > 0; off >>= 1) v = pmin(v, (X)__shfl_xor(v, off, T));
  return v;
}

template <typename T>
__device__ void solve_island_team(const Dev<T>& d, int w, int bstart, int nbi, int cstart, int nci, int last_allow,
                                  int period, char* lds, int t, int TS, SolveMarks& mk) {
  const IslandK c = island_k(d, w, bstart);
  double* sx = (double*)lds;        // [3 * nbi]
  int* cnt = (int*)(sx + 3 * nbi);  // [nbi]
  // 1 + 2. the integrated velocities and the constraint records come from k_prep_bodies and
  // k_prep_contacts; the velocities go to LDS
  for (int i = t; i < nbi; i += TS) {
    const PV<T> p = d.pv[bstart + i];
    sx[3 * i] = p.vx;
    sx[3 * i + 1] = p.vy;
    sx[3 * i + 2] = p.w;
    cnt[i] = 0;
  }
  // 3 + 4. warm start and velocity sweeps as one dataflow sequence
  const int s0 = c.warm ? 0 : 1;
  const int nsw = d.vel_iters + 1 - s0;
  const int mine = t < nci ? (nci - 1 - t) / TS + 1 : 0;  // contacts owned by this lane
  // Lane t owns entries t, t + TS, ... of the island's level order (i_order).  Each of its
  // records learns the lane's entry after it (nxt), so a lane reads its next entry with the
  // record it solves and no pass waits on a separate read of the order.  Only this lane reads
  // these records later, and its reads follow its own writes.
  const uint32_t* order = d.i_order + cstart;
  for (int p = t; p < nci; p += TS) {
    const uint32_t e = order[p], n = p + TS < nci ? order[p + TS] : 0u;
    d.vc[cstart + (e & 0xFFFFu)].nxt = n;
    d.pc[cstart + (e & 0xFFFFu)].nxt = n;
  }
  team_sync();
  // The lane takes its solves (sweep s, owned entry j) in order of s * period + level
  // (sweep_period): a merge of two cursors, A on sweep sa and B on sweep sa + 1.
  {
    const uint32_t e0 = mine > 0 ? order[t] : 0u;
    bool a_on = nsw > 0 && mine > 0, b_on = nsw > 1 && mine > 0;
    int sa = 0, ja = 0, jb = 0;  // each cursor's next entry (ea, eb)
    uint32_t ea = e0, eb = e0;
    int took = 0;  // the cursor whose entry the record just solved was: 1 A, 2 B
    // the lane's next solve: B when its key (sa + 1) * period + level is below A's; nxt is the
    // entry after the one just solved (the solved record's nxt)
    auto next_solve = [&](int& s_out, int& k_out, uint32_t nxt) -> bool {
      // selects only: a branch per case costs more than the few operations it skips
      const bool ta = took == 1, tb = took == 2;
      const bool a_done = ta & (ja >= mine);  // A finished sweep sa (B is never done first)
      const bool shift = a_done & b_on;       // B becomes A, B restarts
      a_on = a_on & !(a_done & !b_on);
      const uint32_t eb1 = tb ? nxt : eb;  // (unused once jb reaches mine)
      ea = shift ? eb1 : (ta ? nxt : ea);
      ja = shift ? jb : ja;
      eb = shift ? e0 : eb1;
      jb = shift ? 0 : jb;
      sa += shift ? 1 : 0;
      b_on = shift ? sa + 1 < nsw : b_on;
      const bool take_b = b_on & (jb < mine) & ((int)(eb >> 16) + period < (int)(ea >> 16));
      s_out = sa + (take_b ? 1 : 0);
      k_out = (int)((take_b ? eb : ea) & 0xFFFFu);
      jb += take_b ? 1 : 0;
      ja += take_b ? 0 : 1;
      took = take_b ? 2 : 1;
      return a_on;
    };
    // A lane holds one record in registers, read when it advances: a second, prefetched
    // record would have to be copied into place at every advance, which costs more than the
    // read latency it hides (a lane rarely takes two solves in consecutive passes).
    int s = 0, k = 0;
    VC<T> cur;
    bool active = next_solve(s, k, 0u);
    if (active) cur = d.vc[cstart + k];
    // The impulses a solve stores go out after the next record's read is issued (a read issued
    // after a store would wait for the store to complete), from registers of their own that
    // stay live into the next pass.  When the next solve is the same contact (its next sweep),
    // the read returns the impulses from before the store: `fresh` has the solve take them from
    // those registers instead.
    ImpPair st0{0.0, 0.0}, st1{0.0, 0.0};
    bool fresh = false;
    const int cap = 2 * nsw * nci + 64;
    // The pass test sits at the bottom and readiness is computed without branches, so the
    // compiler keeps the record where it is (inactive lanes and static sides read entry 0 and
    // ignore it).
    int pass = 0;
    for (bool more = wave_any(active); more; more = wave_any(active && pass < cap)) {
      ++pass;
      asm volatile("" ::"v"(st0.ni), "v"(st0.ti), "v"(st1.ni), "v"(st1.ti));
      {
        const int ia = cur.ia, ib = cur.ib;
        const int ra = s * (int)((unsigned)cur.rka >> 16) + (cur.rka & 0xFFFF);
        const int rb = s * (int)((unsigned)cur.rkb >> 16) + (cur.rkb & 0xFFFF);
        const int xa = active && ia >= 0 ? ia : 0, xb = active && ib >= 0 ? ib : 0;
        // The bodies' velocities are read with their counters: when a counter shows the
        // contact's turn, the velocity it pairs with was written before the last pass boundary.
        const int na = lds_ld(&cnt[xa]), nb_ = lds_ld(&cnt[xb]);
        const double pax = sx[3 * xa], pay = sx[3 * xa + 1], paw = sx[3 * xa + 2];
        const double pbx = sx[3 * xb], pby = sx[3 * xb + 1], pbw = sx[3 * xb + 2];
        const bool ready = active & ((ia < 0) | (na == ra)) & ((ib < 0) | (nb_ == rb));
        if (ready) {
          // static bodies have zero velocity (checked at upload)
          V2d va = v2<double>(ia >= 0 ? pax : 0.0, ia >= 0 ? pay : 0.0);
          V2d vb = v2<double>(ib >= 0 ? pbx : 0.0, ib >= 0 ? pby : 0.0);
          double wa = ia >= 0 ? paw : 0.0, wb = ib >= 0 ? pbw : 0.0;
          cur.p[0].ni = fresh ? st0.ni : cur.p[0].ni;
          cur.p[0].ti = fresh ? st0.ti : cur.p[0].ti;
          cur.p[1].ni = fresh ? st1.ni : cur.p[1].ni;
          cur.p[1].ti = fresh ? st1.ti : cur.p[1].ti;
          const bool solved = s + s0 != 0;
          if (solved) {
            solve_contact_velocity(cur, va, wa, vb, wb);
          } else {
            warm_contact(cur, va, wa, vb, wb);
          }
          if (ia >= 0) { sx[3 * ia] = va.x; sx[3 * ia + 1] = va.y; sx[3 * ia + 2] = wa; lds_st(&cnt[ia], ra + 1); }
          if (ib >= 0) { sx[3 * ib] = vb.x; sx[3 * ib + 1] = vb.y; sx[3 * ib + 2] = wb; lds_st(&cnt[ib], rb + 1); }
          // (a one-point contact stores its unused second point unchanged)
          st0 = ImpPair{cur.p[0].ni, cur.p[0].ti};
          st1 = ImpPair{cur.p[1].ni, cur.p[1].ti};
          asm volatile("" : "+v"(st0.ni), "+v"(st0.ti), "+v"(st1.ni), "+v"(st1.ti));
          ImpPair* dst = reinterpret_cast<ImpPair*>(&d.vc[cstart + k].p[0].ni);
          const int k_done = k;
          active = next_solve(s, k, cur.nxt);
          fresh = k == k_done;
          cur = d.vc[cstart + k];  // (a lane that is done reads its last record again)
          if (solved) {
            dst[0] = st0;
            dst[sizeof(VCP<T>) / sizeof(ImpPair)] = st1;
          }
        }
      }
      pass_fence();
    }
    if (active) atomicOr(&d.w_err[w], WERR_SCHEDULE);  // pass limit reached
  }
  team_sync();
  if (d.i_time) mk.t_vel = (long long)wall_clock64();
  // 5. store impulses, 6. integrate positions (positions to LDS)
  for (int k = t; k < nci; k += TS) store_impulse(d, c, cstart, k);
  for (int i = t; i < nbi; i += TS) {
    PV<T> p = d.pv[bstart + i];
    p.vx = sx[3 * i];
    p.vy = sx[3 * i + 1];
    p.w = sx[3 * i + 2];
    integrate_position(p, c.dt);
    d.pv[bstart + i] = p;
    sx[3 * i] = p.cx;
    sx[3 * i + 1] = p.cy;
    sx[3 * i + 2] = p.a;
  }
  // 7. position iterations with early exit, one dataflow sweep each
  bool position_solved = false;
  for (int it = 0; it < d.pos_iters; ++it) {
    if (!position_solved) {  // team-uniform
      if (d.i_time) mk.npos = it + 1;
      for (int i = t; i < nbi; i += TS) cnt[i] = 0;
      team_sync();
      T min_sep = 0;
      bool active = mine > 0;
      int j = t;
      PC<T> cur;  // one record per lane, read when it advances (as in the velocity sweeps)
      if (active) cur = d.pc[cstart + (order[j] & 0xFFFFu)];
      int pass = 0;
      const int pcap = 2 * nci + 64;
      for (bool more = wave_any(active); more; more = wave_any(active && pass < pcap)) {
        ++pass;
        // the record's ranks and next entry are taken here, a pass after its read: derived
        // right after the read (where the compiler would move them), they would wait for it
        asm volatile("" : "+v"(cur.rka), "+v"(cur.rkb), "+v"(cur.nxt));
        // readiness without branches (a branchy test makes the compiler copy the record at
        // every join); inactive lanes and static sides read entry 0 and ignore it
        const int ia = cur.ia, ib = cur.ib;
        const int ra = cur.rka & 0xFFFF, rb = cur.rkb & 0xFFFF;
        const int xa = active && ia >= 0 ? ia : 0, xb = active && ib >= 0 ? ib : 0;
        const int na = lds_ld(&cnt[xa]), nb_ = lds_ld(&cnt[xb]);
        const double pax = sx[3 * xa], pay = sx[3 * xa + 1], paa = sx[3 * xa + 2];  // read with the counters
        const double pbx = sx[3 * xb], pby = sx[3 * xb + 1], pbb = sx[3 * xb + 2];
        const bool adv = active & ((ia < 0) | (na == ra)) & ((ib < 0) | (nb_ == rb));
        if (adv) {
          double ax = ia >= 0 ? pax : cur.scx, ay = ia >= 0 ? pay : cur.scy, aa = ia >= 0 ? paa : cur.sa;
          double bx = ib >= 0 ? pbx : cur.scx, by = ib >= 0 ? pby : cur.scy, bb = ib >= 0 ? pbb : cur.sa;
          solve_contact_position(cur, ax, ay, aa, bx, by, bb, min_sep);
          if (ia >= 0) { sx[3 * ia] = ax; sx[3 * ia + 1] = ay; sx[3 * ia + 2] = aa; lds_st(&cnt[ia], ra + 1); }
          if (ib >= 0) { sx[3 * ib] = bx; sx[3 * ib + 1] = by; sx[3 * ib + 2] = bb; lds_st(&cnt[ib], rb + 1); }
        }
        if (adv) {
          j += TS;
          active = j < nci;
          if (active) cur = d.pc[cstart + (cur.nxt & 0xFFFFu)];
        }
        pass_fence();
      }
      if (active) atomicOr(&d.w_err[w], WERR_SCHEDULE);
      min_sep = team_min(min_sep, TS);  // the sequential loop's min_separation
      team_sync();
      if (min_sep >= -(T)(3.0 * LINEAR_SLOP)) position_solved = true;
    }
  }
  // 8. copy back
  for (int i = t; i < nbi; i += TS) {
    PV<T> p = d.pv[bstart + i];
    p.cx = sx[3 * i];
    p.cy = sx[3 * i + 1];
    p.a = sx[3 * i + 2];
    write_back(d, c, bstart, i, p);
  }
  // 9. sleep (each lane reads back only the bodies it wrote)
  if (c.wflags & WF_ALLOW_SLEEP) {
    double min_sleep = 1.0e300;
    for (int i = t; i < nbi; i += TS) min_sleep = pmin(min_sleep, sleep_body(d, c, bstart, i, last_allow));
    min_sleep = team_min(min_sleep, TS);
    if (min_sleep >= TIME_TO_SLEEP && position_solved && nbi > 0 && t == (nbi - 1) % TS)
      sleep_last(d, c, bstart, nbi);
  }
  team_sync();  // the next islands reuse the LDS
}

// 1. integrate velocities (island.py:71-95) of every island body, and 2. the constraint records
// of every island contact (contact.py:376-496), as flat launches ahead of k_solve: the work
// has no dependencies inside an island, so it need not sit at the head of each island's
// solve.  k_island_prefix lays the worlds' island lists end to end (exclusive prefixes of
// w_itot); the two kernels then stride over the flat index, each block finding the world of
// its first entry by binary search and its threads stepping on from there, so no block is
// launched for a world's unused capacity.
constexpr int kPrefixBlock = 1024;
constexpr int kPrepGrid = 4096;  // blocks of the flat setup kernels (each strides over the entries)
template <typename T>
__global__ void __launch_bounds__(kPrefixBlock) k_island_prefix(Dev<T> d) {
  // both prefixes in one pass: each thread sums a contiguous run of worlds, the runs' sums are
  // scanned in LDS, and each thread writes its run's prefixes
  __shared__ int sb[kPrefixBlock], sc[kPrefixBlock];
  const int nw = d.nw, t = threadIdx.x;
  const int per = (nw + kPrefixBlock - 1) / kPrefixBlock;
  const int beg = min(nw, t * per), end = min(nw, beg + per);
  int xb = 0, xc = 0;
  for (int w = beg; w < end; ++w) {
    const int2 v = d.w_itot[w];
    xb += v.x;
    xc += v.y;
  }
  sb[t] = xb;
  sc[t] = xc;
  __syncthreads();
  for (int off = 1; off < kPrefixBlock; off <<= 1) {
    const int vb = t >= off ? sb[t - off] : 0, vc = t >= off ? sc[t - off] : 0;
    __syncthreads();
    sb[t] += vb;
    sc[t] += vc;
    __syncthreads();
  }
  int* pb = d.w_ipre;
  int* pc = d.w_ipre + nw + 1;
  int rb = sb[t] - xb, rc = sc[t] - xc;
  for (int w = beg; w < end; ++w) {
    const int2 v = d.w_itot[w];
    pb[w] = rb;
    pc[w] = rc;
    rb += v.x;
    rc += v.y;
  }
  if (t == kPrefixBlock - 1) {
    const int tb = sb[t], tc = sc[t];
    pb[nw] = tb;
    pc[nw] = tc;
    // the island bodies and contacts of the step, for the work counters
    if (tb > 0) atomicAdd(&d.stat[S_IBODIES], (unsigned long long)tb);
    if (tc > 0) atomicAdd(&d.stat[S_ICONTACTS], (unsigned long long)tc);
  }
}

// The world of flat entry i: w with pre[w] <= i < pre[w + 1], from w_first (the world of the
// block's first entry) on; worlds without island entries have pre[w] == pre[w + 1].
__device__ __forceinline__ int world_of(const int* pre, int w_first, int i) {
  int w = w_first;
  while (pre[w + 1] <= i) ++w;
  return w;
}
// The last world w with pre[w] <= i (block-uniform, by thread 0 into LDS).
__device__ __forceinline__ int first_world(const int* pre, int nw, int i) {
  int lo = 0, hi = nw;  // pre[lo] <= i < pre[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (pre[mid] <= i) lo = mid; else hi = mid;
  }
  return lo;
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_prep_bodies(Dev<T> d) {
  __shared__ int w_first;
  const int* pre = d.w_ipre;
  const int total = pre[d.nw];
  for (int base = blockIdx.x * kBlock; base < total; base += gridDim.x * kBlock) {
    if (threadIdx.x == 0) w_first = first_world(pre, d.nw, base);
    __syncthreads();
    const int i = base + threadIdx.x;
    if (i < total) {
      const int w = world_of(pre, w_first, i);
      const int b0 = d.w_body_off[w];
      const IslandK c = island_k(d, w, b0);
      const int k = i - pre[w];
      d.pv[b0 + k] = integrate_velocity(d, c, b0, k);
    }
    __syncthreads();
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_prep_contacts(Dev<T> d) {
  __shared__ int w_first;
  const int* pre = d.w_ipre + d.nw + 1;
  const int total = pre[d.nw];
  for (int base = blockIdx.x * kBlock; base < total; base += gridDim.x * kBlock) {
    if (threadIdx.x == 0) w_first = first_world(pre, d.nw, base);
    __syncthreads();
    const int i = base + threadIdx.x;
    if (i < total) {
      const int w = world_of(pre, w_first, i);
      const int k = i - pre[w];
      const int b0 = d.w_body_off[w], cb = d.w_con_base[w];
      const int bstart = b0 + d.i_isb[cb + k];  // the contact's island
      const IslandK c = island_k(d, w, bstart);
      // each contact's rank in its bodies' sequences and their degrees: rank | degree << 16
      const uint2 r = d.i_rank[cb + k];
      const int rka = (int)((r.x & 0xFFFFu) | (r.y & 0xFFFFu) << 16);
      const int rkb = (int)((r.x >> 16) | (r.y >> 16) << 16);
      int ia, ib;
      init_contact(d, c, bstart, cb, k, ia, ib, rka, rkb);
    }
    __syncthreads();
  }
}

// Island timing (b2g_island_times): wall-clock ticks around one island's solve, its size and
// the lanes that solved it (a team's size, 1 for the lane path, 0 for a team island solved by
// one lane because its state did not fit the team's LDS share).
template <typename T>
__device__ __forceinline__ void record_island_time(const Dev<T>& d, int q, long long t0, int nci, int nbi, int lanes,
                                                   int period, const SolveMarks& mk) {
  long long* o = d.i_time + kIslandTimeFields * (size_t)q;
  o[0] = t0;
  o[1] = (long long)wall_clock64();
  o[2] = (long long)nci | (long long)nbi << 32;
  o[3] = (long long)lanes | (long long)period << 32;
  o[4] = mk.t_vel;
  o[5] = mk.npos;
}

// One wavefront per block.  The queue is ordered largest class first.  Its
// entries of classes >= heavy_class are dealt to wavefronts 64 / T islands at
// a time (T = team_size of the class), each team with its share of the LDS
// budget (queue_class has moved an island whose state does not fit its class's
// share to the top class, a whole wavefront and its budget; one that does not
// fit even that is solved by the team's first lane alone).  The remaining
// islands are dealt 64 per wavefront, one lane each.  Islands with joints are
// left to k_solve_joints.  Wavefronts are taken in queue order, so the longest
// islands start first and are spread over the whole chip.
// Two wavefronts per SIMD at least: the FP64 records would otherwise take
// more than 256 VGPRs and halve the resident wavefronts.
template <typename T>
__global__ void __launch_bounds__(kSolveBlock) __attribute__((amdgpu_waves_per_eu(2))) k_solve(Dev<T> d) {
  extern __shared__ __align__(16) char team_lds[];
  const int n = *d.q_count;
  const int nh = d.q_cls[kQHeavy];
  const int hw = d.q_cls[kQHeavyWaves];
  const int nwaves = hw + (n - nh + kSolveBlock - 1) / kSolveBlock;
  const int l = threadIdx.x;
  for (int g = blockIdx.x; g < nwaves; g += gridDim.x) {
    if (g < hw) {
      // classes own consecutive wave ranges, highest class first
      int cls = kSizeClasses - 1;
      while (cls > d.heavy_class && g >= d.q_cls[kQWaveBase + cls - 1]) --cls;
      const int TS = team_size(cls, d.team_full_class, d.team_max);
      const int per = kSolveBlock / TS;
      const int team = l / TS, t = l % TS;
      const int q = d.q_cls[kQBase + cls] + (g - d.q_cls[kQWaveBase + cls]) * per + team;
      const int qend = d.q_cls[kQBase + cls] + d.q_cls[cls];
      if (q < qend) {
        const int4 in = d.qs_isl[q];
        const int nbi = in.z & 0xFFFF, period = (int)((uint32_t)in.z >> 16);
        const int nc = d.qs_nc[q];
        const int nci = nc & 0x3FFFFFFF, last_allow = (nc >> 30) & 1;
        const int share = d.team_lds / per;
        const long long t0 = d.i_time ? (long long)wall_clock64() : 0;
        const bool fits = team_lds_bytes(nbi) <= share;
        SolveMarks mk{0, 0};
        if (island_has_joints(d, in.x, in.y)) {
          // k_solve_joints takes it
        } else if (fits) {
          solve_island_team(d, in.x, in.y, nbi, in.w, nci, last_allow, period, team_lds + team * share, t, TS, mk);
          if (d.i_time && t == 0) record_island_time(d, q, t0, nci, nbi, TS, period, mk);
        } else if (t == 0) {
          solve_island(d, in.x, in.y, nbi, in.w, nci, last_allow, mk);
          if (d.i_time) record_island_time(d, q, t0, nci, nbi, 0, period, mk);
        }
      }
      team_sync();
    } else {
      const int q = nh + (g - hw) * kSolveBlock + l;
      if (q < n) {
        const long long t0 = d.i_time ? (long long)wall_clock64() : 0;
        const int4 in = d.qs_isl[q];
        const int nc = d.qs_nc[q];
        SolveMarks mk{0, 0};
        if (!island_has_joints(d, in.x, in.y)) {  // else k_solve_joints takes it
          solve_island(d, in.x, in.y, in.z & 0xFFFF, in.w, nc & 0x3FFFFFFF, (nc >> 30) & 1, mk);
          if (d.i_time)
            record_island_time(d, q, t0, nc & 0x3FFFFFFF, in.z & 0xFFFF, 1, (int)((uint32_t)in.z >> 16), mk);
        }
      }
    }
  }
}

// The islands with joints, one lane each, after k_solve (which leaves them): launched only for
// batches with joints.  Their bodies and contacts belong to no other island, so the order of the
// two kernels does not matter.
template <typename T>
__global__ void __launch_bounds__(64) k_solve_joints(Dev<T> d) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= *d.q_count) return;
  const int4 in = d.qs_isl[q];
  if (!island_has_joints(d, in.x, in.y)) return;
  const long long t0 = d.i_time ? (long long)wall_clock64() : 0;
  const int nc = d.qs_nc[q];
  const int nbi = in.z & 0xFFFF, nci = nc & 0x3FFFFFFF;
  SolveMarks mk{0, 0};
  solve_island_joints(d, in.x, in.y, nbi, in.w, nci, (nc >> 30) & 1, mk);
  if (d.i_time) record_island_time(d, q, t0, nci, nbi, 0, (int)((uint32_t)in.z >> 16), mk);
}

// ---------------------------------------------------------------------------
// fixture synchronisation + move_proxy (FP64)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void compute_aabb64(int type, int n, const double* g, double r, double px, double py,
                                               double c, double s, double& lx, double& ly, double& ux, double& uy) {
  if (type == SHAPE_CIRCLE) {  // shapes.py:178-189: p = xf.p + R * position
    double qx = c * g[0] + (-s) * g[1], qy = s * g[0] + c * g[1];
    double x = px + qx, y = py + qy;
    lx = x - r; ly = y - r; ux = x + r; uy = y + r;
  } else if (type == SHAPE_POLYGON) {  // shapes.py:316-334
    double x0 = px + c * g[0] + (-s) * g[1], y0 = py + s * g[0] + c * g[1];
    double lox = x0, loy = y0, hix = x0, hiy = y0;
    for (int i = 0; i < n; ++i) {
      double x = px + c * g[2 * i] + (-s) * g[2 * i + 1];
      double y = py + s * g[2 * i] + c * g[2 * i + 1];
      lox = pmin(lox, x); loy = pmin(loy, y);
      hix = pmax(hix, x); hiy = pmax(hiy, y);
    }
    lx = lox - r; ly = loy - r; ux = hix + r; uy = hiy + r;
  } else {  // edge, shapes.py:702-718
    double ax = px + c * g[0] + (-s) * g[1], ay = py + s * g[0] + c * g[1];
    double bx = px + c * g[2] + (-s) * g[3], by = py + s * g[2] + c * g[3];
    lx = pmin(ax, bx) - r; ly = pmin(ay, by) - r; ux = pmax(ax, bx) + r; uy = pmax(ay, by) + r;
  }
}

template <typename T>
__device__ __forceinline__ bool sync_fixture(const Dev<T>& d, int f);

template <typename T>
__global__ void __launch_bounds__(kBlock) k_sync(Dev<T> d) {
  const int f = blockIdx.x * kBlock + threadIdx.x;
  const bool moved = f < d.nf && sync_fixture(d, f);
  const unsigned long long m = __ballot(moved);  // every lane of the wavefront gets here
  if ((threadIdx.x & 63) == 0 && m) atomicAdd(&d.stat[S_MOVED], (unsigned long long)__popcll(m));
}

// Fixture._synchronize (fixture.py:287-298) of fixture f on body gb, with DynamicTree.move_proxy
// (dynamictree.py:152-188): returns whether its proxy was re-fattened.
template <typename T>
__device__ __forceinline__ bool sync_fixture_aabb(const Dev<T>& d, int f, int gb) {
  int4 info = d.f_info[f];
  const int type = info.y & 255, n = (info.y >> 8) & 255;
  const double* g = d.f_geo64 + (size_t)f * 16;
  const double r = d.f_rad64[f];
  double2 c0 = d.b_c0[gb], lc = d.b_lc[gb];
  double s0, k0;
  sincos(d.b_a0[gb], &s0, &k0);
  double x1 = c0.x - (k0 * lc.x + (-s0) * lc.y), y1 = c0.y - (s0 * lc.x + k0 * lc.y);
  double4 xf = d.b_xf[gb];
  double alx, aly, aux, auy, blx, bly, bux, buy;
  compute_aabb64(type, n, g, r, x1, y1, k0, s0, alx, aly, aux, auy);
  compute_aabb64(type, n, g, r, xf.x, xf.y, xf.z, xf.w, blx, bly, bux, buy);
  double lx = pmin(alx, blx), ly = pmin(aly, bly), ux = pmax(aux, bux), uy = pmax(auy, buy);
  double dx = xf.x - x1, dy = xf.y - y1;
  double4 fat = d.f_fat[f];
  if (fat.x <= lx && fat.y <= ly && ux <= fat.z && uy <= fat.w) return false;  // contains
  double nlx = lx - AABB_EXTENSION, nly = ly - AABB_EXTENSION, nux = ux + AABB_EXTENSION, nuy = uy + AABB_EXTENSION;
  double mdx = AABB_MULTIPLIER * dx, mdy = AABB_MULTIPLIER * dy;
  if (mdx < 0.0) { nlx = nlx + mdx; nly = nly + 0.0; }
  else { nux = nux + mdx; nuy = nuy + 0.0; }
  if (mdy < 0.0) { nlx = nlx + 0.0; nly = nly + mdy; }
  else { nux = nux + 0.0; nuy = nuy + mdy; }
  d.f_fat[f] = make_double4(nlx, nly, nux, nuy);
  return true;
}

// Body._synchronize_fixtures for fixture f after the solve: returns whether its proxy moved.
template <typename T>
__device__ __forceinline__ bool sync_fixture(const Dev<T>& d, int f) {
  const int gb = d.f_wb[f].y;
  d.f_moved[f] = 0;
  if (d.b_island[gb] < 0 || d.b_type[gb] == BT_STATIC) return false;
  if (!sync_fixture_aabb(d, f, gb)) return false;
  d.f_moved[f] = 1;
  return true;
}

// ---------------------------------------------------------------------------
// broad-phase pairs + contact creation, one workgroup per world
// ---------------------------------------------------------------------------
__host__ __device__ inline int next_pow2(int x) {
  int p = 1;
  while (p < x) p <<= 1;
  return p;
}

__host__ __device__ inline size_t pairs_smem_bytes(int nf, int ccap) {
  size_t s = 0;
  s += ((size_t)nf * 4 + 15) & ~(size_t)15;                 // mpos
  s += ((size_t)nf * 2 + 15) & ~(size_t)15;                 // mlist
  s += (size_t)next_pow2(ccap + 1) * 4;                     // hash
  s += (size_t)next_pow2(8 * nf + 64) * 4;                  // new pairs
  s += (size_t)kBigPairsBlock * 4 + 64;  // scan scratch for the widest workgroup
  return s;
}

__device__ __forceinline__ uint32_t hkey(int a, int b) {
  return a < b ? ((uint32_t)a << 16) | (uint32_t)b : ((uint32_t)b << 16) | (uint32_t)a;
}
__device__ __forceinline__ uint32_t hhash(uint32_t k) {
  k ^= k >> 16;
  k *= 0x7feb352du;
  k ^= k >> 15;
  k *= 0x846ca68bu;
  k ^= k >> 16;
  return k;
}

template <typename T>
__device__ void copy_contact(const Dev<T>& d, int dst, int src) {
  d.c_pair[dst] = d.c_pair[src];
  d.c_state[dst] = d.c_state[src];
  d.c_mat[dst] = d.c_mat[src];
  d.c_lnlp[dst] = d.c_lnlp[src];
  d.c_pts[dst] = d.c_pts[src];
  d.c_imp[dst] = d.c_imp[src];
  d.c_id[dst] = d.c_id[src];
}

// Contact._create (contact.py:313-326) + Contact.__init__ (contact.py:55-90):
// the registry orientation, mixed friction and restitution, an empty manifold.
template <typename T>
__device__ __forceinline__ void create_contact(const Dev<T>& d, int f0, int s, int P, int R) {
  const int ta = d.f_info[f0 + P].y & 255, tb = d.f_info[f0 + R].y & 255;
  const bool swap = (ta == SHAPE_CIRCLE && tb == SHAPE_POLYGON) || (ta == SHAPE_CIRCLE && tb == SHAPE_EDGE) ||
                    (ta == SHAPE_POLYGON && tb == SHAPE_EDGE);
  const int fa = swap ? R : P, fb = swap ? P : R;
  d.c_pair[s] = make_int2(fa, fb);
  d.c_state[s] = CS_LIVE | (M_CIRCLES << CS_MTYPE_SHIFT) | (swap ? CS_SWAPPED : 0);
  V2<T> ma = d.f_mat[f0 + fa], mb = d.f_mat[f0 + fb];
  d.c_mat[s] = v2<T>(sqrt(ma.x * mb.x), pmax(ma.y, mb.y));
  d.c_lnlp[s] = V4<T>{0, 0, 0, 0};
  d.c_pts[s] = V4<T>{0, 0, 0, 0};
  d.c_imp[s] = V4<double>{0, 0, 0, 0};
  d.c_id[s] = make_uint2(0, 0);
}

__device__ __forceinline__ double readlane_f64(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, l), hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// The move buffer as BroadPhase.update_pairs walks it (broadphase.py:61-68, 86-99): first the
// entries the batch brought that no search has taken yet (proxies created, or touched by
// Fixture.refilter), in their order; after a solve, then the proxies the synchronisation moved,
// in fixture order (bodies in order, fixtures body-major).  A proxy keeps its first place: a later
// entry of it adds no pair (the pair set).  mpos[f] = place of fixture f or -1, mlist[place] = f.
// list_batch_moves (every thread of the block) lists the batch's entries and returns their count.
template <int B, typename T, typename IdxT>
__device__ int list_batch_moves(const Dev<T>& d, int w, int f0, int nf, int* mpos, IdxT* mlist) {
  for (int i = threadIdx.x; i < nf; i += B) mpos[i] = -1;
  __syncthreads();
  const int nup = d.w_move_n[w];
  if (threadIdx.x == 0)
    for (int k = 0; k < nup; ++k) {
      const int f = d.mv[f0 + k];  // distinct and in range (checked at upload)
      mpos[f] = k;
      mlist[k] = (IdxT)f;
    }
  __syncthreads();
  return nup;
}

// The synchronisation's moved proxies not yet listed, after the nup listed ones; f_moved (set by
// k_sync in this step, read by nothing after the pair search) serves as the scan's array.
// Returns the list's length.
template <int B, typename T, typename IdxT>
__device__ int append_synced_moves(const Dev<T>& d, int f0, int nf, int* mpos, IdxT* mlist, int nup, int* scan) {
  int* fm = d.f_moved + f0;
  for (int i = threadIdx.x; i < nf; i += B) fm[i] = (fm[i] && mpos[i] < 0) ? 1 : 0;
  __syncthreads();
  const int total = block_exclusive_scan<B>(fm, nf, scan);
  for (int i = threadIdx.x; i < nf; i += B) {
    const int next = i + 1 < nf ? fm[i + 1] : total;
    if (next != fm[i]) {
      mpos[i] = nup + fm[i];
      mlist[nup + fm[i]] = (IdxT)i;
    }
  }
  __syncthreads();
  return nup + total;
}

// A pair of overlapping fat AABBs (moved proxy P at move position i, proxy R): the rest of
// find_new_contacts' filter (contact_manager.py:102-158); a new contact is listed as i << 16 | R.
template <typename T>
__device__ void pair_candidate(const Dev<T>& d, int w, int f0, int b0, int P, int R, int i, const uint32_t* hash,
                               int hcap, uint32_t* nlist, int ncap, int* misc) {
  const int4 ip = d.f_info[f0 + P], ir = d.f_info[f0 + R];
  if (ip.x == ir.x) return;  // same body
  const int tb_p = d.b_type[b0 + ip.x], tb_r = d.b_type[b0 + ir.x];
  if (!(d.b_flags[b0 + ip.x] & BF_ACTIVE) || !(d.b_flags[b0 + ir.x] & BF_ACTIVE)) return;
  {  // exists?
    const uint32_t k = hkey(P, R);
    uint32_t h = hhash(k) & (uint32_t)(hcap - 1);
    for (;;) {
      const uint32_t v = hash[h];
      if (v == 0xFFFFFFFFu) break;
      if (v == k) return;
      h = (h + 1) & (uint32_t)(hcap - 1);
    }
  }
  if (tb_r != BT_DYNAMIC && tb_p != BT_DYNAMIC) return;  // Body._should_collide
  if (joint_blocks(d, w, ip.x, ir.x)) return;
  if (!(d.w_flags[w] & WF_CUSTOM_FILTER)) {  // World.default_contact_filter(fixture P, fixture R)
    bool ok;
    if (ip.w == ir.w && ip.w != 0) ok = ip.w > 0;
    else ok = ((ip.z >> 16) & (ir.z & 0xFFFF)) != 0 && ((ip.z & 0xFFFF) & (ir.z >> 16)) != 0;
    if (!ok) return;
  }
  const int ta = ip.y & 255, tb = ir.y & 255;
  const bool reg = (ta == SHAPE_CIRCLE || ta == SHAPE_POLYGON || ta == SHAPE_EDGE) &&
                   (tb == SHAPE_CIRCLE || tb == SHAPE_POLYGON || tb == SHAPE_EDGE) &&
                   !(ta == SHAPE_EDGE && tb == SHAPE_EDGE);
  if (!reg) return;
  const int slot = atomicAdd(&misc[0], 1);
  if (slot < ncap) nlist[slot] = ((uint32_t)i << 16) | (uint32_t)R;
  else atomicOr(&misc[1], 1);
}

// B threads per workgroup: 256, or 1024 for the largest worlds (those whose LDS footprint
// allows at most two workgroups per CU), whose search sets the kernel's tail.
template <typename T, int B>
__global__ void __launch_bounds__(B) k_pairs(Dev<T> d, int w_first, int from_upload, int finish, int do_pairs) {
  constexpr int kBlock = B;
  extern __shared__ __align__(16) unsigned char smem[];
  const int w = d.w_order[w_first + blockIdx.x], tid = threadIdx.x;  // largest worlds first
  if (d.w_big && d.w_big[w]) return;  // k_pairs_big_*
  const int b0 = d.w_body_off[w], nb = d.w_body_off[w + 1] - b0;
  const int f0 = d.w_fix_off[w], nf = d.w_fix_off[w + 1] - f0;
  const int cb = d.w_con_base[w], ccap = d.w_con_cap[w];
  unsigned char* p = smem;
  int* mpos = (int*)p; p += ((size_t)nf * 4 + 15) & ~(size_t)15;
  uint16_t* mlist = (uint16_t*)p; p += ((size_t)nf * 2 + 15) & ~(size_t)15;
  const int hcap = next_pow2(ccap + 1);
  uint32_t* hash = (uint32_t*)p; p += (size_t)hcap * 4;
  const int ncap = next_pow2(8 * nf + 64);
  uint32_t* nlist = (uint32_t*)p; p += (size_t)ncap * 4;
  int* scan = (int*)p; p += (size_t)kBlock * 4;
  int* misc = (int*)p;

  // ---- the move buffer (broadphase.py:61-68) ----
  int nmove = 0;
  if (do_pairs) {  // dt == 0 without new fixtures: World._solve (and find_new_contacts) does not run
    nmove = list_batch_moves<B>(d, w, f0, nf, mpos, mlist);
    if (!from_upload) nmove = append_synced_moves<B>(d, f0, nf, mpos, mlist, nmove, scan);
  }
  if (tid == 0) { misc[0] = 0; misc[1] = 0; }
  __syncthreads();
  if (nmove > 0) {
    // existing contacts (either orientation), for the "does a contact exist" test
    for (int i = tid; i < hcap; i += kBlock) hash[i] = 0xFFFFFFFFu;
    __syncthreads();
    const int used = d.w_con_used[w];
    for (int sl = tid; sl < used; sl += kBlock) {
      if (!(d.c_state[cb + sl] & CS_LIVE)) continue;
      int2 pr = d.c_pair[cb + sl];
      uint32_t k = hkey(pr.x, pr.y);
      uint32_t h = hhash(k) & (uint32_t)(hcap - 1);
      for (;;) {
        uint32_t prev = atomicCAS(&hash[h], 0xFFFFFFFFu, k);
        if (prev == 0xFFFFFFFFu || prev == k) break;
        h = (h + 1) & (uint32_t)(hcap - 1);
      }
    }
    __syncthreads();
    // every moved proxy against every proxy of the world; exact FP64 tests.  A lane keeps
    // proxy R's fat AABB in registers; the wavefront loads 64 moved proxies at a time, one
    // per lane, and every lane walks them with (uniform) lane reads.
    const int lane = tid & 63;
    for (int base = 0; base < nmove; base += 64) {
      const int nm = nmove - base < 64 ? nmove - base : 64;
      int Pl = 0;
      double4 fl = make_double4(0, 0, 0, 0);
      if (lane < nm) {
        Pl = mlist[base + lane];
        fl = d.f_fat[f0 + Pl];
      }
      for (int R = tid; R < nf; R += kBlock) {
        const double4 fr = d.f_fat[f0 + R];
        const int mr = mpos[R];
        for (int jj = 0; jj < nm; ++jj) {
          const int i = base + jj;
          const int P = __builtin_amdgcn_readlane(Pl, jj);
          if (R == P) continue;
          if (mr >= 0 && mr < i) continue;  // (R, P) is already in the pair set
          const double4 fp = make_double4(readlane_f64(fl.x, jj), readlane_f64(fl.y, jj), readlane_f64(fl.z, jj),
                                          readlane_f64(fl.w, jj));
          if (!aabb_overlap(fr.x, fr.y, fr.z, fr.w, fp.x, fp.y, fp.z, fp.w)) continue;
          pair_candidate(d, w, f0, b0, P, R, i, hash, hcap, nlist, ncap, misc);
        }
      }
    }
    __syncthreads();
  }
  int nnew = misc[0];
  if (misc[1]) {
    if (tid == 0) atomicOr(&d.w_err[w], WERR_NEWPAIRS);
    nnew = 0;
  }
  if (nnew > 0) {
    // canonical order: move-buffer position, then partner index (bitonic sort)
    const int np = next_pow2(nnew);
    for (int i = nnew + tid; i < np; i += kBlock) nlist[i] = 0xFFFFFFFFu;
    __syncthreads();
    for (int k = 2; k <= np; k <<= 1) {
      for (int j = k >> 1; j > 0; j >>= 1) {
        for (int i = tid; i < np; i += kBlock) {
          int ixj = i ^ j;
          if (ixj > i) {
            uint32_t a = nlist[i], b = nlist[ixj];
            bool up = (i & k) == 0;
            if ((a > b) == up) { nlist[i] = b; nlist[ixj] = a; }
          }
        }
        __syncthreads();
      }
    }
    // make room: squeeze out destroyed slots when the segment would overflow
    int used = d.w_con_used[w];
    if (used + nnew > ccap) {
      int base = 0;
      for (int c0 = 0; c0 < used; c0 += kBlock) {
        const int sl = c0 + tid;
        const bool live = sl < used && (d.c_state[cb + sl] & CS_LIVE);
        int2 pr = make_int2(0, 0);
        int st = 0;
        V2<T> mat = v2<T>(0, 0);
        V4<T> lnlp{}, pts{};
        V4<double> imp{};
        uint2 id = make_uint2(0, 0);
        if (live) {
          pr = d.c_pair[cb + sl]; st = d.c_state[cb + sl]; mat = d.c_mat[cb + sl];
          lnlp = d.c_lnlp[cb + sl]; pts = d.c_pts[cb + sl]; imp = d.c_imp[cb + sl]; id = d.c_id[cb + sl];
        }
        scan[tid] = live ? 1 : 0;
        __syncthreads();
        for (int off = 1; off < kBlock; off <<= 1) {
          int v = tid >= off ? scan[tid - off] : 0;
          __syncthreads();
          scan[tid] += v;
          __syncthreads();
        }
        const int dst = base + scan[tid] - (live ? 1 : 0);
        const int cnt = scan[kBlock - 1];
        __syncthreads();
        if (live) {
          const int s = cb + dst;
          d.c_pair[s] = pr; d.c_state[s] = st; d.c_mat[s] = mat; d.c_lnlp[s] = lnlp; d.c_pts[s] = pts;
          d.c_imp[s] = imp; d.c_id[s] = id;
        }
        base += cnt;
        __syncthreads();
      }
      used = base;
    }
    if (used + nnew > ccap) {
      if (tid == 0) atomicOr(&d.w_err[w], WERR_CAPACITY);
      nnew = 0;
    }
    for (int k = tid; k < nnew; k += kBlock) {
      const uint32_t key = nlist[k];
      create_contact(d, f0, cb + used + k, (int)mlist[key >> 16], (int)(key & 0xFFFFu));
    }
    __syncthreads();
    if (tid == 0) {
      d.w_con_used[w] = used + nnew;
      if (nnew > 0) atomicAdd(&d.stat[S_NEW], (unsigned long long)nnew);
    }
  }
  if (tid == 0) {
    if (do_pairs) d.w_move_n[w] = 0;  // update_pairs empties the buffer (broadphase.py:101-102)
    d.w_new[w] = nnew;
  }
  if (!finish) return;
  // ---- end-of-step bookkeeping (world.py:348-352) ----
  if (tid == 0) {
    if (d.dt > 0.0) d.w_inv_dt0[w] = d.inv_dt;
    misc[2] = 0;
  }
  __syncthreads();
  const bool clear = !(d.w_flags[w] & WF_KEEP_FORCES);
  int asleep = 0;
  for (int i = tid; i < nb; i += kBlock) {
    const int gb = b0 + i;
    if (clear) {
      d.b_f[gb] = v2<T>(0, 0);
      d.b_t[gb] = (T)0;
    }
    if (d.b_type[gb] != BT_STATIC && !(d.b_flags[gb] & BF_AWAKE)) ++asleep;
  }
  if (asleep) atomicAdd(&misc[2], asleep);
  __syncthreads();
  if (tid == 0) {
    d.w_asleep[w] = misc[2];
    if (d.w_flags[w] & WF_NEW_FIXTURE) d.w_flags[w] &= ~WF_NEW_FIXTURE;
  }
}

// ---------------------------------------------------------------------------
// Worlds too large for one workgroup's LDS (e.g. one world of 11 200 proxies)
//
// The island search and the pair search of such worlds keep their state in
// global memory.  The awake gating becomes a fixed point over "the slot of the
// contact that woke the body" (a contact is evaluated when one of its bodies
// was awake before its slot, as in the in-order loop); the islands are the
// union-find components, over touching contacts between non-static bodies,
// that hold an awake body; each island's depth-first search then runs on its
// own lane from the island's first awake body, which visits bodies and
// contacts in exactly the order the world-wide search would.  New pairs come
// from a many-workgroup brute-force FP64 overlap test, are sorted into the
// canonical order and created by one workgroup per world.
// ---------------------------------------------------------------------------
constexpr int kBigBlock = 1024;
constexpr size_t kBigSortLds = 65536;  // k_islands_big: at least 512 contacts per wavefront's level sort
constexpr size_t kBigWalkLdsMax = 147456;  // k_islands_big: most dynamic LDS for the walks' state
constexpr int kIntMax = 0x7fffffff;

__device__ __forceinline__ int g_ld(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// State of the large-world per-island walks: body state bits, rank counters (rank | first
// level << 16), the level of each body's last contact, and the contact-seen flags; in global
// memory, or in LDS when the world's fits (then one global round trip per adjacency entry,
// which the walk prefetches ahead, is all that is left of the walk's memory latency).
struct WalkGlobal {
  int* bst; int* cur; int* lev; int* cflag;
  __device__ int bst_ld(int i) const { return g_ld(&bst[i]); }
  __device__ void bst_st(int i, int v) const { bst[i] = v; }
  __device__ int cur_ld(int i) const { return g_ld(&cur[i]); }
  __device__ void cur_st(int i, int v) const { cur[i] = v; }
  __device__ int lev_ld(int i) const { return g_ld(&lev[i]); }
  __device__ void lev_st(int i, int v) const { lev[i] = v; }
  __device__ int cf_ld(int sl) const { return g_ld(&cflag[sl]); }
  __device__ void cf_set(int sl) const { cflag[sl] = 1; }
};
struct WalkLds {
  int* bst; int* cur; uint16_t* lev; uint32_t* cf;
  __device__ int bst_ld(int i) const { return bst[i]; }
  __device__ void bst_st(int i, int v) const { bst[i] = v; }
  __device__ int cur_ld(int i) const { return cur[i]; }
  __device__ void cur_st(int i, int v) const { cur[i] = v; }
  __device__ int lev_ld(int i) const { return lev[i]; }
  __device__ void lev_st(int i, int v) const { lev[i] = (uint16_t)v; }
  __device__ int cf_ld(int sl) const { return (int)((cf[sl >> 5] >> (sl & 31)) & 1u); }
  __device__ void cf_set(int sl) const { atomicOr(&cf[sl >> 5], 1u << (sl & 31)); }  // words shared by islands
};
__host__ __device__ inline size_t walk_lds_bytes(int nb, int ccap) {
  return ((size_t)nb * 4 + 15) / 16 * 16 * 2 + ((size_t)nb * 2 + 15) / 16 * 16 + ((size_t)(ccap + 31) / 32 * 4 + 15) / 16 * 16;
}
// Union-find over a large world's bodies (lock-free hooking of the larger root under the
// smaller, path halving) on the world's parent array in global memory, shared by the
// workgroups of k_islands_big_ends (device-scope atomics: the XCDs' L2s are not coherent for
// plain accesses).  The unions end with that kernel; k_islands_big then finds the roots
// read-only.
__device__ __forceinline__ int gu_get(int* P, int x) {
  return __hip_atomic_load(&P[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ int gu_find(int* P, int x) {
  for (;;) {
    const int p = gu_get(P, x);
    if (p == x) return x;
    const int gp = gu_get(P, p);
    if (gp != p) atomicCAS(&P[x], p, gp);  // path halving (skipped if x moved meanwhile)
    x = gp;
  }
}
__device__ void gu_unite(int* P, int a, int b) {
  for (;;) {
    a = gu_find(P, a);
    b = gu_find(P, b);
    if (a == b) return;
    if (a < b) { const int t = a; a = b; b = t; }
    if (atomicCAS(&P[a], a, b) == a) return;  // the root with the larger index joins the smaller
  }
}
__device__ __forceinline__ int gu_root(const int* P, int x) {
  for (int p = P[x]; p != x; p = P[x]) x = p;
  return x;
}

// The touching contacts among kEndsBatch slots base, base + stride, ... of a large world, with
// their bodies (world-local).  The three dependent global reads (state, fixture pair, fixture
// bodies) are issued for all slots of the batch before any is waited for: one lane of the
// single workgroup otherwise waits out each chain in turn.
constexpr int kEndsBatch = 4;
template <typename T>
__device__ __forceinline__ void touching_ends(const Dev<T>& d, int cb, int f0, int used, int base, bool* ok,
                                              int* la, int* lb, int stride = kBigBlock) {
  int st[kEndsBatch];
#pragma unroll
  for (int u = 0; u < kEndsBatch; ++u) {
    const int sl = base + u * stride;
    st[u] = sl < used ? d.c_state[cb + sl] : 0;
  }
  int2 pr[kEndsBatch];
#pragma unroll
  for (int u = 0; u < kEndsBatch; ++u) {
    ok[u] = (st[u] & (CS_LIVE | CS_TOUCHING | CS_ENABLED)) == (CS_LIVE | CS_TOUCHING | CS_ENABLED);
    pr[u] = ok[u] ? d.c_pair[cb + base + u * stride] : make_int2(0, 0);
  }
#pragma unroll
  for (int u = 0; u < kEndsBatch; ++u) {
    la[u] = ok[u] ? d.f_info[f0 + pr[u].x].x : 0;
    lb[u] = ok[u] ? d.f_info[f0 + pr[u].y].x : 0;
  }
}

// LDS written by some lanes of a wavefront is read by others: keep the accesses in order.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// The large-world island search, in nine launches: k_islands_big_head (one workgroup per
// world) sets up the per-body state and runs the awake gating; k_islands_big_ends (many
// workgroups per world) takes the touching contacts' ends, the bodies' degrees and the
// union-find over them; k_islands_big_scan (one) the adjacency offsets; k_islands_big_fill
// and k_islands_big_sort (many) the adjacency in slot order and the component roots;
// k_islands_big_seeds and k_islands_big_counts (many) the islands' seeds and sizes; and
// k_islands_big (one workgroup per world) the islands' ranges, the walks, the level sorts and
// the queue; and k_islands_big_tail (many) the island contacts' records and the bodies' flags.
template <typename T>
__global__ void __launch_bounds__(kBigBlock) k_islands_big_head(Dev<T> d) {
  __shared__ int sh[8];
  const int w = d.big_list[blockIdx.x], tid = threadIdx.x;
  const int b0 = d.w_body_off[w], nb = d.w_body_off[w + 1] - b0;
  const int f0 = d.w_fix_off[w];
  const int cb = d.w_con_base[w];
  const int used = d.w_con_used[w];
  int* bst = d.g_bst + b0;
  int* wake = d.g_wake + b0;
  int* off = d.g_off + b0 + w;
  int* lev = d.g_lev + b0;
  int* parent = d.g_parent + b0;
  int* seed = d.g_seed + b0;
  int* cntb = d.g_cntb + b0;
  int* cntc = d.g_cntc + b0;
  int* cflag = d.g_cflag + cb;
  for (int i = tid; i < nb; i += kBigBlock) {
    const int fl = d.b_flags[b0 + i], ty = d.b_type[b0 + i];
    bst[i] = (ty == BT_STATIC ? 1 : 0) | ((fl & BF_AWAKE) ? 2 : 0) | ((fl & BF_ALLOW_SLEEP) ? 4 : 0) |
             ((fl & BF_ACTIVE) ? 8 : 0);
    off[i] = 0;
    lev[i] = 0;
    parent[i] = i;
    seed[i] = kIntMax;
    cntb[i] = 0;
    cntc[i] = 0;
  }
  for (int sl = tid; sl < used; sl += kBigBlock) cflag[sl] = 0;
  if (tid < 8) sh[tid] = 0;
  __syncthreads();

  // --- awake gating (contact_manager.py:191-197) as a fixed point over wake slots ---
  if (d.w_asleep[w] > 0 && !d.solve_only) {
    for (int i = tid; i < nb; i += kBigBlock) wake[i] = (bst[i] & 2) ? -1 : kIntMax;
    __syncthreads();
    for (;;) {
      if (tid == 0) sh[0] = 0;
      __syncthreads();
      for (int sl = tid; sl < used; sl += kBigBlock) {
        const int s = cb + sl;
        int st = d.c_state[s];
        if (!(st & CS_LIVE)) continue;
        const int2 pr = d.c_pair[s];
        const int la = d.f_info[f0 + pr.x].x, lb = d.f_info[f0 + pr.y].x;
        if (!(st & CS_PROCESSED)) {
          const bool act = (!(bst[la] & 1) && g_ld(&wake[la]) < sl) || (!(bst[lb] & 1) && g_ld(&wake[lb]) < sl);
          if (!act) continue;
          st = narrow_update(d, w, s, st);
          d.c_state[s] = st;
          if (!(st & CS_LIVE)) continue;
        }
        if (st & CS_WOKE) {
          if (atomicMin(&wake[la], sl) > sl) sh[0] = 1;
          if (atomicMin(&wake[lb], sl) > sl) sh[0] = 1;
        }
      }
      __syncthreads();
      const int changed = sh[0];
      __syncthreads();
      if (!changed) break;
    }
    for (int i = tid; i < nb; i += kBigBlock) {
      const int wk = wake[i];
      if (wk >= 0 && wk < kIntMax && !(bst[i] & 2)) bst[i] |= 2 | 32;
    }
    __syncthreads();
  }

}

// The touching-contact graph's first pass: each touching contact adds to its bodies' degrees
// and unites them (two non-static bodies); grid (kBigEndsGrid, big worlds) x kBlock.
constexpr int kBigEndsGrid = 64;
template <typename T>
__global__ void __launch_bounds__(kBlock) k_islands_big_ends(Dev<T> d) {
  const int w = d.big_list[blockIdx.y];
  const int b0 = d.w_body_off[w], f0 = d.w_fix_off[w], cb = d.w_con_base[w], used = d.w_con_used[w];
  const int* bst = d.g_bst + b0;
  int* off = d.g_off + b0 + w;
  int* parent = d.g_parent + b0;
  const int stride = gridDim.x * kBlock;
  for (int base = blockIdx.x * kBlock + threadIdx.x; base < used; base += kEndsBatch * stride) {
    bool ok[kEndsBatch];
    int la[kEndsBatch], lb[kEndsBatch];
    touching_ends(d, cb, f0, used, base, ok, la, lb, stride);
    int sa[kEndsBatch], sb[kEndsBatch];
#pragma unroll
    for (int u = 0; u < kEndsBatch; ++u) {
      sa[u] = ok[u] ? bst[la[u]] : 1;
      sb[u] = ok[u] ? bst[lb[u]] : 1;
    }
#pragma unroll
    for (int u = 0; u < kEndsBatch; ++u) {
      if (!ok[u]) continue;
      atomicAdd(&off[la[u]], 1);
      atomicAdd(&off[lb[u]], 1);
      if (!(sa[u] & 1) && !(sb[u] & 1)) gu_unite(parent, la[u], lb[u]);
    }
  }
}

// The bodies' adjacency offsets (exclusive scan of the degrees), one workgroup per world.
template <typename T>
__global__ void __launch_bounds__(kBigBlock) k_islands_big_scan(Dev<T> d) {
  __shared__ int scan[kBigBlock];
  const int w = d.big_list[blockIdx.x], tid = threadIdx.x;
  const int b0 = d.w_body_off[w], nb = d.w_body_off[w + 1] - b0;
  int* off = d.g_off + b0 + w;
  int* cur = d.g_cur + b0;
  const int total = block_exclusive_scan<kBigBlock>(off, nb, scan);
  if (tid == 0) off[nb] = total;
  for (int i = tid; i < nb; i += kBigBlock) cur[i] = off[i];
}

// The adjacency entries (contact slot, other body) of every touching contact at its two bodies;
// grid (kBigEndsGrid, big worlds) x kBlock.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_islands_big_fill(Dev<T> d) {
  const int w = d.big_list[blockIdx.y];
  const int b0 = d.w_body_off[w], f0 = d.w_fix_off[w], cb = d.w_con_base[w], used = d.w_con_used[w];
  int* cur = d.g_cur + b0;
  int2* adj = d.g_adj + 2 * (size_t)cb;
  const int stride = gridDim.x * kBlock;
  for (int base = blockIdx.x * kBlock + threadIdx.x; base < used; base += kEndsBatch * stride) {
    bool ok[kEndsBatch];
    int la[kEndsBatch], lb[kEndsBatch];
    touching_ends(d, cb, f0, used, base, ok, la, lb, stride);
    int pa[kEndsBatch], pb[kEndsBatch];
#pragma unroll
    for (int u = 0; u < kEndsBatch; ++u) {
      pa[u] = ok[u] ? atomicAdd(&cur[la[u]], 1) : 0;
      pb[u] = ok[u] ? atomicAdd(&cur[lb[u]], 1) : 0;
    }
#pragma unroll
    for (int u = 0; u < kEndsBatch; ++u) {
      if (!ok[u]) continue;
      const int sl = base + u * stride;
      adj[pa[u]] = make_int2(sl, lb[u]);
      adj[pb[u]] = make_int2(sl, la[u]);
    }
  }
}

// Each body's entries in ascending slot order (= body._contacts order) and, for a non-static
// body, its component's root into the wake slots (free again); grid (kBigEndsGrid, big worlds)
// x kBlock.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_islands_big_sort(Dev<T> d) {
  const int w = d.big_list[blockIdx.y];
  const int b0 = d.w_body_off[w], nb = d.w_body_off[w + 1] - b0, cb = d.w_con_base[w];
  const int* bst = d.g_bst + b0;
  const int* off = d.g_off + b0 + w;
  const int* parent = d.g_parent + b0;
  int* cur = d.g_cur + b0;
  int* wake = d.g_wake + b0;
  int2* adj = d.g_adj + 2 * (size_t)cb;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < nb; i += gridDim.x * kBlock) {
    cur[i] = 0;  // from here on: contacts of the body seen by the island search
    const int beg = off[i], end = off[i + 1];
    for (int k = beg + 1; k < end; ++k) {
      const int2 e = adj[k];
      int j = k - 1;
      while (j >= beg && adj[j].x > e.x) {
        adj[j + 1] = adj[j];
        --j;
      }
      adj[j + 1] = e;
    }
    if (!(bst[i] & 1)) wake[i] = gu_root(parent, i);
  }
}

// Islands are the components with an awake, active, non-static body; seed = the first such
// body.  Each thread takes a contiguous run of bodies (then of slots) and adds to a
// component's seed and counts once per run of equal roots: the bodies and contacts of one
// component mostly sit together, and one atomic per body on a few hot roots serialises in L2.
// Grid (kBigEndsGrid, big worlds) x kBlock.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_islands_big_seeds(Dev<T> d) {
  const int w = d.big_list[blockIdx.y];
  const int b0 = d.w_body_off[w], nb = d.w_body_off[w + 1] - b0;
  const int* bst = d.g_bst + b0;
  const int* wake = d.g_wake + b0;  // the roots (k_islands_big_sort)
  int* parent = d.g_parent + b0;
  int* seed = d.g_seed + b0;
  const int nt = gridDim.x * kBlock, t = blockIdx.x * kBlock + threadIdx.x;
  const int bper = (nb + nt - 1) / nt, bbeg = min(nb, t * bper), bend = min(nb, bbeg + bper);
  int root = -1, first = kIntMax;
  for (int i = bbeg; i < bend; ++i) {
    const int bs = bst[i];
    if (bs & 1) continue;
    const int r = wake[i];
    parent[i] = r;
    if (!(bs & 2) || !(bs & 8)) continue;
    if (r != root) {
      if (root >= 0) atomicMin(&seed[root], first);
      root = r;
      first = i;
    }
  }
  if (root >= 0) atomicMin(&seed[root], first);
}

// Each island's bodies and contacts (components with a seed); grid (kBigEndsGrid, big worlds)
// x kBlock.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_islands_big_counts(Dev<T> d) {
  const int w = d.big_list[blockIdx.y];
  const int b0 = d.w_body_off[w], nb = d.w_body_off[w + 1] - b0;
  const int f0 = d.w_fix_off[w], cb = d.w_con_base[w], used = d.w_con_used[w];
  const int* bst = d.g_bst + b0;
  const int* parent = d.g_parent + b0;
  const int* seed = d.g_seed + b0;
  int* cntb = d.g_cntb + b0;
  int* cntc = d.g_cntc + b0;
  const int nt = gridDim.x * kBlock, t = blockIdx.x * kBlock + threadIdx.x;
  {
    const int bper = (nb + nt - 1) / nt, bbeg = min(nb, t * bper), bend = min(nb, bbeg + bper);
    int root = -1, n = 0;
    for (int i = bbeg; i < bend; ++i) {
      if (bst[i] & 1) continue;
      const int r = parent[i];
      if (r != root) {
        if (root >= 0 && n > 0) atomicAdd(&cntb[root], n);
        root = r;
        n = seed[r] != kIntMax ? 1 : 0;
      } else if (n > 0) {
        ++n;
      }
    }
    if (root >= 0 && n > 0) atomicAdd(&cntb[root], n);
  }
  {
    const int sper = (used + nt - 1) / nt, sbeg = min(used, t * sper), send = min(used, sbeg + sper);
    int root = -1, n = 0;
    for (int base = sbeg; base < send; base += kEndsBatch) {
      bool ok[kEndsBatch];
      int la[kEndsBatch], lb[kEndsBatch];
      touching_ends(d, cb, f0, send, base, ok, la, lb, 1);
      int x[kEndsBatch];
#pragma unroll
      for (int u = 0; u < kEndsBatch; ++u) x[u] = ok[u] && (bst[la[u]] & 1) ? lb[u] : la[u];
      int px[kEndsBatch];
#pragma unroll
      for (int u = 0; u < kEndsBatch; ++u) px[u] = ok[u] ? parent[x[u]] : 0;
      int sd[kEndsBatch];
#pragma unroll
      for (int u = 0; u < kEndsBatch; ++u) sd[u] = ok[u] ? seed[px[u]] : kIntMax;
#pragma unroll
      for (int u = 0; u < kEndsBatch; ++u) {
        if (!ok[u] || sd[u] == kIntMax) continue;
        if (px[u] != root) {
          if (root >= 0) atomicAdd(&cntc[root], n);
          root = px[u];
          n = 0;
        }
        ++n;
      }
    }
    if (root >= 0 && n > 0) atomicAdd(&cntc[root], n);
  }
}

template <typename T>
__global__ void __launch_bounds__(kBigBlock) k_islands_big(Dev<T> d, int sort_cap, int smem_dyn) {
  __shared__ int scan[kBigBlock];
  __shared__ int sh[8];
  // each wavefront's 2 * sort_cap words for the level sort of islands of up to sort_cap
  // contacts, or the walks' state
  extern __shared__ uint32_t uf_lds[];
  const int w = d.big_list[blockIdx.x], tid = threadIdx.x;
  const int b0 = d.w_body_off[w], nb = d.w_body_off[w + 1] - b0;
  const int f0 = d.w_fix_off[w];
  const int cb = d.w_con_base[w], ccap = d.w_con_cap[w];
  const int used = d.w_con_used[w];
  int* bst = d.g_bst + b0;
  int* wake = d.g_wake + b0;
  int* off = d.g_off + b0 + w;
  int* cur = d.g_cur + b0;
  int* lev = d.g_lev + b0;
  int* parent = d.g_parent + b0;
  int* seed = d.g_seed + b0;
  int* cntb = d.g_cntb + b0;
  int* cntc = d.g_cntc + b0;
  int* bidx = d.g_bidx + b0;
  int2* adj = d.g_adj + 2 * (size_t)cb;
  int* cflag = d.g_cflag + cb;
  int4* isl = d.g_isl + b0;
  int* isl_b = d.i_bodies + b0;
  int* isl_c = d.i_contacts + cb;
  if (tid < 8) sh[tid] = 0;
  __syncthreads();
  // --- islands: components with an awake, active, non-static body; seed = the first such body
  // (k_islands_big_seeds and k_islands_big_counts) ---
  unsigned long long* alloc = (unsigned long long*)(d.g_wmisc + 16 * w);
  if (tid == 0) *alloc = 0ull;
  __syncthreads();
  for (int i = tid; i < nb; i += kBigBlock) {
    if ((bst[i] & 1) || parent[i] != i || seed[i] == kIntMax) continue;
    const unsigned long long r = atomicAdd(alloc, ((unsigned long long)cntb[i] << 32) | (unsigned)cntc[i]);
    const int k = atomicAdd(&sh[1], 1);
    isl[k] = make_int4(i, (int)(r >> 32), (int)(r & 0xffffffffu), seed[i]);
  }
  __syncthreads();
  const int nisl = sh[1];
  const int nbi_tot = (int)(*alloc >> 32), nci_tot = (int)(*alloc & 0xffffffffu);
  // --- depth-first search of each island (world.py:494-580), one lane per island ---
  const bool lds_walk = walk_lds_bytes(nb, ccap) <= (size_t)smem_dyn;
  WalkLds LW{};
  if (lds_walk) {  // the union-find parents are dead: the walk state moves into LDS
    unsigned char* p = (unsigned char*)uf_lds;
    LW.bst = (int*)p; p += ((size_t)nb * 4 + 15) / 16 * 16;
    LW.cur = (int*)p; p += ((size_t)nb * 4 + 15) / 16 * 16;
    LW.lev = (uint16_t*)p; p += ((size_t)nb * 2 + 15) / 16 * 16;
    LW.cf = (uint32_t*)p;
    for (int i = tid; i < nb; i += kBigBlock) {
      LW.bst[i] = bst[i];
      LW.cur[i] = 0;
      LW.lev[i] = 0;
    }
    for (int i = tid; i < (ccap + 31) / 32; i += kBigBlock) LW.cf[i] = 0u;
    __syncthreads();
  }
  auto walk = [&](const auto& S) {
  for (int k = tid; k < nisl; k += kBigBlock) {
    const int4 r = isl[k];
    const int bstart = r.y, cstart = r.z, s0 = r.w;
    const int cap_b = cntb[r.x], cap_c = cntc[r.x];
    int* stack = d.g_stack + b0 + 2 * (size_t)cb + bstart + 2 * cstart;  // cap_b + 2 cap_c entries
    int* statics = stack + cap_b + 2 * cap_c;                              // grows downwards
    int nstat = 0, sp = 0, nbi = 0, nci = 0, last_allow = 0, span = 0, depth = 0;
    stack[sp++] = s0;
    S.bst_st(s0, S.bst_ld(s0) | 16);
    bool bad = false;
    while (sp > 0 && !bad) {
      const int b = stack[--sp];
      int bs = S.bst_ld(b);
      last_allow = (bs >> 2) & 1;
      if (!(bs & 2)) bs |= 2 | 32;  // body.awake = True
      S.bst_st(b, bs);
      if (bs & 1) continue;  // static bodies join but do not propagate
      if (nbi >= cap_b) { bad = true; break; }
      bidx[b] = bstart + nbi;  // position in the world's island-body list, like the LDS search
      wake[b] = bstart;        // the start of its island (the wake slots are free again)
      isl_b[bstart + nbi++] = b;
      // The expanding body's counters stay in registers (no other lane touches this island);
      // cur = rank | first level << 16.  Adjacency entries (read only) are read four steps
      // ahead, and an entry's contact flag and other body's state are read together (relaxed
      // atomic loads in global memory, which the compiler keeps ahead of the test).
      int cur_b = S.cur_ld(b), lev_b = S.lev_ld(b);
      const int ebeg = off[b], eend = off[b + 1];
      int2 a0 = ebeg < eend ? adj[ebeg] : make_int2(0, 0);
      int2 a1 = ebeg + 1 < eend ? adj[ebeg + 1] : make_int2(0, 0);
      int2 a2 = ebeg + 2 < eend ? adj[ebeg + 2] : make_int2(0, 0);
      int2 a3 = ebeg + 3 < eend ? adj[ebeg + 3] : make_int2(0, 0);
      for (int e = ebeg; e < eend; ++e) {
        const int2 a4 = e + 4 < eend ? adj[e + 4] : make_int2(0, 0);
        const int sl_ = a0.x, o = a0.y;
        a0 = a1;
        a1 = a2;
        a2 = a3;
        a3 = a4;
        const int fw = S.cf_ld(sl_);
        const int so = S.bst_ld(o);
        const int co = S.cur_ld(o);
        const int lo_read = S.lev_ld(o);
        if (fw) continue;
        if (nci >= cap_c || sp + nstat + 2 > cap_b + 2 * cap_c) { bad = true; break; }
        S.cf_set(sl_);
        const bool o_dyn = !(so & 1);
        const int lo = o_dyn ? lo_read : 0;
        const int lvl = 1 + (lev_b > lo ? lev_b : lo);
        const uint32_t rank_b = (uint32_t)(cur_b & 0xFFFF);
        const int fb = rank_b == 0 ? lvl : (int)((uint32_t)cur_b >> 16);
        cur_b = (int)(rank_b + 1) | fb << 16;
        span = max(span, lvl - fb);
        uint32_t rank_o = 0u;
        if (o_dyn) {
          rank_o = (uint32_t)(co & 0xFFFF);
          const int fo = rank_o == 0 ? lvl : (int)((uint32_t)co >> 16);
          S.cur_st(o, (int)(rank_o + 1) | fo << 16);
          span = max(span, lvl - fo);
          S.lev_st(o, lvl);
        }
        depth = max(depth, lvl);
        lev_b = lvl;
        d.i_rank[cb + cstart + nci] = make_uint2((uint32_t)b, rank_b | rank_o << 16);
        d.i_order[cb + cstart + nci] = (uint32_t)(lvl - 1);  // level, until sorted below
        isl_c[cstart + nci++] = sl_;
        if (o_dyn) {
          if (so & 16) continue;
          S.bst_st(o, so | 16);
        } else {
          bool seen = false;
          for (int q = 1; q <= nstat; ++q) seen = seen || statics[-q] == o;
          if (seen) continue;
          statics[-(++nstat)] = o;
        }
        stack[sp++] = o;
      }
      S.cur_st(b, cur_b);
      S.lev_st(b, lev_b);
    }
    // Island-local contact indices and level keys are 16-bit (i_order, the solver's ranks):
    // an island of more than 65535 contacts is reported and not solved.
    const bool too_big = nci > 0xFFFF;
    if (bad || nbi != cap_b || nci != cap_c || too_big) {  // the component sizes and the search disagree
      atomicOr(&d.w_err[w], too_big && !bad && nbi == cap_b && nci == cap_c ? WERR_ISLAND_SIZE : WERR_SCHEDULE);
      for (int q = nci; q < cap_c; ++q) isl_c[cstart + q] = 0;  // keep later passes in range
      for (int q = nbi; q < cap_b; ++q) isl_b[bstart + q] = s0;
      nbi = nci = 0;
    }
    d.i_info[b0 + k] = make_int4(bstart, nbi | sweep_period(span, depth) << 16, cstart, nci | (last_allow << 30));
    if (nci <= sort_cap) continue;  // sorted by a wavefront in LDS below
    // the island's contacts by level (counting sort in the stack area)
    int* hist = stack;
    for (int q = 0; q < nci; ++q) hist[q] = 0;
    for (int q = 0; q < nci; ++q) hist[d.i_order[cb + cstart + q]] += 1;
    int run = 0;
    for (int q = 0; q < nci; ++q) {
      const int v = hist[q];
      hist[q] = run;
      run += v;
    }
    for (int q = 0; q < nci; ++q) {
      const int lv = (int)d.i_order[cb + cstart + q];
      stack[nci + hist[lv]] = q | lv << 16;  // island-local index | level key << 16
      hist[lv] += 1;
    }
    for (int q = 0; q < nci; ++q) d.i_order[cb + cstart + q] = (uint32_t)stack[nci + q];
  }
  };
  if (lds_walk) {
    walk(LW);
    __syncthreads();
    for (int i = tid; i < nb; i += kBigBlock) {  // what the later phases read
      bst[i] = LW.bst[i];
      cur[i] = LW.cur[i];
    }
  } else {
    walk(WalkGlobal{bst, cur, lev, cflag});
  }
  __syncthreads();
  // Islands of up to sort_cap contacts: their contacts by level, one wavefront per island, a
  // counting sort in LDS (levels 0 .. nci - 1).  Contacts of one level share no dynamic body,
  // so their order within the level does not matter.
  {
    const int lane = tid & 63, wv = tid >> 6;
    int* h = (int*)uf_lds + (size_t)wv * 2 * sort_cap;
    int* o = h + sort_cap;
    for (int k = wv; k < nisl; k += kBigBlock / 64) {
      const int4 info = d.i_info[b0 + k];
      const int nci = info.w & 0x3FFFFFFF;
      if (nci == 0 || nci > sort_cap) continue;
      uint32_t* ord = d.i_order + cb + info.z;
      for (int q = lane; q < nci; q += 64) h[q] = 0;
      wave_lds_sync();
      for (int q = lane; q < nci; q += 64) atomicAdd(&h[ord[q]], 1);
      wave_lds_sync();
      const int per = (nci + 63) / 64, lo = min(nci, lane * per), hi = min(nci, lo + per);
      int sum = 0;
      for (int i = lo; i < hi; ++i) sum += h[i];
      int incl = sum;
      for (int sh_ = 1; sh_ < 64; sh_ <<= 1) {
        const int v = __shfl_up(incl, sh_, 64);
        if (lane >= sh_) incl += v;
      }
      int run = incl - sum;
      for (int i = lo; i < hi; ++i) {
        const int v = h[i];
        h[i] = run;
        run += v;
      }
      wave_lds_sync();
      for (int q = lane; q < nci; q += 64) {
        const int lv = (int)ord[q];
        o[atomicAdd(&h[lv], 1)] = q | lv << 16;  // island-local index | level key << 16
      }
      wave_lds_sync();
      for (int q = lane; q < nci; q += 64) ord[q] = (uint32_t)o[q];
      wave_lds_sync();
    }
  }
  __syncthreads();
  // --- queue, island-local contact ends, ranks oriented as (body A, body B) ---
  if (tid == 0) sh[2] = nisl > 0 ? atomicAdd(d.q_count, nisl) : 0;
  if (tid < kSizeClasses) scan[tid] = 0;
  __syncthreads();
  const int qbase = sh[2];
  for (int k = tid; k < nisl; k += kBigBlock) {
    const int4 in = d.i_info[b0 + k];
    d.q_isl[qbase + k] = make_int4(w, b0 + in.x, in.y, cb + in.z);
    d.q_nc[qbase + k] = in.w;
    atomicAdd(&scan[queue_class(d, in.w & 0x3FFFFFFF, in.y & 0xFFFF)], 1);
  }
  __syncthreads();
  if (tid < kSizeClasses && scan[tid] > 0) atomicAdd(&d.q_cls[tid], scan[tid]);
  if (tid == 0) {
    d.w_nisl[w] = nisl;
    d.w_itot[w] = make_int2(nbi_tot, nci_tot);
  }
  (void)ccap;
}

// The island contacts' ends (island-local bodies, or -1 - body for a static one), fixtures and
// bodies, island body start, and ranks oriented as (body A, body B) with the bodies' degrees;
// the bodies' woken and island flags.  Grid (kBigEndsGrid, big worlds) x kBlock.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_islands_big_tail(Dev<T> d) {
  const int w = d.big_list[blockIdx.y];
  const int b0 = d.w_body_off[w], nb = d.w_body_off[w + 1] - b0;
  const int f0 = d.w_fix_off[w], cb = d.w_con_base[w];
  const int* bst = d.g_bst + b0;
  const int* wake = d.g_wake + b0;  // island body starts (the walks)
  const int* cur = d.g_cur + b0;
  const int* bidx = d.g_bidx + b0;
  const int* isl_c = d.i_contacts + cb;
  const int nci_tot = (int)(*(const unsigned long long*)(d.g_wmisc + 16 * w) & 0xffffffffu);
  const int nt = gridDim.x * kBlock, t = blockIdx.x * kBlock + threadIdx.x;
  for (int k = t; k < nci_tot; k += nt) {
    const int2 pr = d.c_pair[cb + isl_c[k]];
    const int la = d.f_info[f0 + pr.x].x, lb = d.f_info[f0 + pr.y].x;
    const int ia = (bst[la] & 1) ? -1 - la : bidx[la];
    const int ib = (bst[lb] & 1) ? -1 - lb : bidx[lb];
    d.i_cidx[cb + k] = make_int2(ia, ib);
    d.i_fb[cb + k] = make_uint2((uint32_t)pr.x | (uint32_t)pr.y << 16, (uint32_t)la | (uint32_t)lb << 16);
    d.i_isb[cb + k] = wake[ia >= 0 ? la : lb];
    const uint2 r = d.i_rank[cb + k];
    const uint32_t lo = r.y & 0xFFFFu, hi = r.y >> 16;
    const uint32_t dga = ia >= 0 ? (uint32_t)cur[la] & 0xFFFFu : 0u;  // rank counters end at the degrees
    const uint32_t dgb = ib >= 0 ? (uint32_t)cur[lb] & 0xFFFFu : 0u;
    d.i_rank[cb + k] = make_uint2((int)r.x == la ? (lo | hi << 16) : (hi | lo << 16), dga | dgb << 16);
  }
  for (int i = t; i < nb; i += nt) {
    const int bs = bst[i];
    int fl = d.b_flags[b0 + i];
    if (bs & 32) {  // woken: Body.awake setter (body.py:800-812)
      fl |= BF_AWAKE;
      d.b_sleep[b0 + i] = 0.0;
      d.b_flags[b0 + i] = fl;
    }
    d.b_island[b0 + i] = ((bs & 16) && !(bs & 1)) ? 1 : -1;
  }
}

// Move buffer and the hash of existing contacts of each large world.
template <typename T>
__global__ void __launch_bounds__(kBigBlock) k_pairs_big_prep(Dev<T> d, int from_upload, int do_pairs) {
  __shared__ int scan[kBigBlock];
  const int w = d.big_list[blockIdx.x], tid = threadIdx.x;
  const int f0 = d.w_fix_off[w], nf = d.w_fix_off[w + 1] - f0;
  const int cb = d.w_con_base[w], ccap = d.w_con_cap[w];
  int* mpos = d.g_mpos + f0;
  int* mlist = d.g_mlist + f0;
  int* misc = d.g_wmisc + 16 * w;  // [2] moves, [3] new pairs, [4] overflow
  int nmove = 0;
  if (do_pairs) {  // the move buffer as in k_pairs
    nmove = list_batch_moves<kBigBlock>(d, w, f0, nf, mpos, mlist);
    if (!from_upload) nmove = append_synced_moves<kBigBlock>(d, f0, nf, mpos, mlist, nmove, scan);
  }
  if (tid == 0) {
    misc[2] = nmove;
    misc[3] = 0;
    misc[4] = 0;
  }
  if (nmove > 0) {
    const int hcap = next_pow2(ccap + 1);
    uint32_t* hash = d.g_hash + d.g_hbase[w];
    for (int i = tid; i < hcap; i += kBigBlock) hash[i] = 0xFFFFFFFFu;
    __syncthreads();
    const int used = d.w_con_used[w];
    for (int sl = tid; sl < used; sl += kBigBlock) {
      if (!(d.c_state[cb + sl] & CS_LIVE)) continue;
      const int2 pr = d.c_pair[cb + sl];
      const uint32_t k = hkey(pr.x, pr.y);
      uint32_t h = hhash(k) & (uint32_t)(hcap - 1);
      for (;;) {
        const uint32_t prev = atomicCAS(&hash[h], 0xFFFFFFFFu, k);
        if (prev == 0xFFFFFFFFu || prev == k) break;
        h = (h + 1) & (uint32_t)(hcap - 1);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// LBVH over a world's fat AABBs (the tree the reference keeps in DynamicTree,
// dynamictree.py:115-207, built from scratch instead of updated incrementally):
// 30-bit Morton codes of the proxies' centres, a bitonic sort of (code, proxy)
// keys in LDS (in global scratch for worlds over kTreeLdsKeys proxies), Karras'
// radix-tree construction from the sorted keys (the proxy index in the low bits
// makes every key distinct), and a bottom-up refit in sweeps.  The tree's shape is
// not observable: queries and the pair search test the proxies' own fat AABBs
// exactly, and internal nodes only prune (their boxes are exact unions).
// ---------------------------------------------------------------------------
constexpr int kTreeEmpty = 0x7FFFFFFF;
constexpr int kTreeLdsKeys = 16384;  // 128 KB of LDS
constexpr int kTreeBlock = 256;
constexpr int kTreeStack = 64;       // 62-bit distinct keys: at most 62 levels

__device__ __forceinline__ uint32_t morton_spread15(uint32_t v) {
  v &= 0x7FFFu;
  v = (v | (v << 8)) & 0x00FF00FFu;
  v = (v | (v << 4)) & 0x0F0F0F0Fu;
  v = (v | (v << 2)) & 0x33333333u;
  v = (v | (v << 1)) & 0x55555555u;
  return v;
}

__device__ __forceinline__ int tree_delta(const unsigned long long* k, int n, int i, int j) {
  if (j < 0 || j >= n) return -1;
  return __clzll(k[i] ^ k[j]);
}

// In-place bitonic sort of keys[0, np) (np a power of two) by one workgroup.
__device__ void block_bitonic_sort(unsigned long long* keys, int np) {
  for (int size = 2; size <= np; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int t = threadIdx.x; t < (np >> 1); t += blockDim.x) {
        const int lo = 2 * t - (t & (stride - 1)), hi = lo + stride;
        const bool up = (lo & size) == 0;
        const unsigned long long a = keys[lo], b = keys[hi];
        if ((a > b) == up) {
          keys[lo] = b;
          keys[hi] = a;
        }
      }
      __syncthreads();
    }
  }
}

template <typename T>
__global__ void __launch_bounds__(kTreeBlock) k_tree_build(Dev<T> d, const int* wlist) {
  extern __shared__ unsigned long long tkeys[];
  __shared__ double red[4][kTreeBlock / 64];
  __shared__ int sweep_done;
  const int w = wlist[blockIdx.x], tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int f0 = d.w_fix_off[w], n = d.w_fix_off[w + 1] - f0;
  if (n < 2) {
    if (tid == 0) d.t_root[w] = n == 1 ? ~0 : kTreeEmpty;
    if (n == 1 && tid == 0) d.t_key[f0] = 0ull;
    return;
  }
  // bounds of the proxies' centres
  double mnx = 1e300, mny = 1e300, mxx = -1e300, mxy = -1e300;
  for (int i = tid; i < n; i += kTreeBlock) {
    const double4 b = d.f_fat[f0 + i];
    const double cx = 0.5 * (b.x + b.z), cy = 0.5 * (b.y + b.w);
    mnx = fmin(mnx, cx); mny = fmin(mny, cy); mxx = fmax(mxx, cx); mxy = fmax(mxy, cy);
  }
  for (int o = 32; o > 0; o >>= 1) {
    mnx = fmin(mnx, __shfl_xor(mnx, o, 64)); mny = fmin(mny, __shfl_xor(mny, o, 64));
    mxx = fmax(mxx, __shfl_xor(mxx, o, 64)); mxy = fmax(mxy, __shfl_xor(mxy, o, 64));
  }
  if (lane == 0) { red[0][wv] = mnx; red[1][wv] = mny; red[2][wv] = mxx; red[3][wv] = mxy; }
  __syncthreads();
  mnx = red[0][0]; mny = red[1][0]; mxx = red[2][0]; mxy = red[3][0];
  for (int k = 1; k < kTreeBlock / 64; ++k) {
    mnx = fmin(mnx, red[0][k]); mny = fmin(mny, red[1][k]); mxx = fmax(mxx, red[2][k]); mxy = fmax(mxy, red[3][k]);
  }
  const double sx = mxx > mnx ? 32767.0 / (mxx - mnx) : 0.0, sy = mxy > mny ? 32767.0 / (mxy - mny) : 0.0;
  // keys
  int np = 1;
  while (np < n) np <<= 1;
  const bool in_lds = np <= kTreeLdsKeys;
  unsigned long long* keys = in_lds ? tkeys : d.t_sort + d.t_sbase[w];
  for (int i = tid; i < np; i += kTreeBlock) {
    unsigned long long key = ~0ull;
    if (i < n) {
      const double4 b = d.f_fat[f0 + i];
      const double cx = 0.5 * (b.x + b.z), cy = 0.5 * (b.y + b.w);
      const uint32_t qx = (uint32_t)fmin(32767.0, fmax(0.0, (cx - mnx) * sx));
      const uint32_t qy = (uint32_t)fmin(32767.0, fmax(0.0, (cy - mny) * sy));
      const uint32_t code = morton_spread15(qx) | (morton_spread15(qy) << 1);
      key = ((unsigned long long)code << 32) | (unsigned long long)i;
    }
    keys[i] = key;
  }
  __syncthreads();
  block_bitonic_sort(keys, np);
  for (int i = tid; i < n; i += kTreeBlock) d.t_key[f0 + i] = keys[i];
  // hierarchy (Karras 2012): internal node i covers the sorted range [min(i, j), max(i, j)]
  for (int i = tid; i < n - 1; i += kTreeBlock) {
    const int dir = tree_delta(keys, n, i, i + 1) - tree_delta(keys, n, i, i - 1) >= 0 ? 1 : -1;
    const int dmin = tree_delta(keys, n, i, i - dir);
    int lmax = 2;
    while (tree_delta(keys, n, i, i + lmax * dir) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
      if (tree_delta(keys, n, i, i + (l + t) * dir) > dmin) l += t;
    const int j = i + l * dir;
    const int dnode = tree_delta(keys, n, i, j);
    int sp = 0;
    for (int t = (l + 1) >> 1;; t = (t + 1) >> 1) {
      if (tree_delta(keys, n, i, i + (sp + t) * dir) > dnode) sp += t;
      if (t == 1) break;
    }
    const int g = i + sp * dir + (dir < 0 ? -1 : 0);
    const int left = (i < j ? i : j) == g ? ~g : g;
    const int right = (i > j ? i : j) == g + 1 ? ~(g + 1) : g + 1;
    d.t_child[f0 + i] = make_int2(left, right);
    d.t_flag[f0 + i] = 0;
  }
  __syncthreads();
  // refit in sweeps: a node whose children are done takes the union of their boxes (phase 1,
  // flag -1), and becomes done after the barrier (phase 2), until the root is done
  for (;;) {
    for (int i = tid; i < n - 1; i += kTreeBlock) {
      if (d.t_flag[f0 + i] != 0) continue;
      const int2 c = d.t_child[f0 + i];
      if ((c.x >= 0 && d.t_flag[f0 + c.x] != 1) || (c.y >= 0 && d.t_flag[f0 + c.y] != 1)) continue;
      double4 a, b;
      if (c.x < 0) a = d.f_fat[f0 + (int)(keys[~c.x] & 0xFFFFFFFFu)];
      else a = d.t_box[f0 + c.x];
      if (c.y < 0) b = d.f_fat[f0 + (int)(keys[~c.y] & 0xFFFFFFFFu)];
      else b = d.t_box[f0 + c.y];
      d.t_box[f0 + i] = make_double4(fmin(a.x, b.x), fmin(a.y, b.y), fmax(a.z, b.z), fmax(a.w, b.w));
      d.t_flag[f0 + i] = -1;
    }
    __syncthreads();
    for (int i = tid; i < n - 1; i += kTreeBlock)
      if (d.t_flag[f0 + i] == -1) d.t_flag[f0 + i] = 1;
    __syncthreads();
    if (tid == 0) sweep_done = d.t_flag[f0] == 1;
    __syncthreads();
    if (sweep_done) break;
    __syncthreads();
  }
  if (tid == 0) d.t_root[w] = 0;
}

// Visits every proxy of world w whose subtree boxes overlap the box (inclusive), with the leaf's
// world-local fixture index; the visitor applies the exact test on the proxy's own fat AABB.
// Returns false if the stack overflowed (a tree deeper than its keys allow: never).
template <typename T, typename Visit>
__device__ __forceinline__ bool tree_visit(const Dev<T>& d, int w, double lx, double ly, double ux, double uy,
                                           Visit&& visit) {
  const int f0 = d.w_fix_off[w];
  const int root = d.t_root[w];
  if (root == kTreeEmpty) return true;
  int stack[kTreeStack];
  int sp = 0;
  stack[sp++] = root;
  while (sp > 0) {
    const int node = stack[--sp];
    if (node < 0) {
      visit((int)(d.t_key[f0 + ~node] & 0xFFFFFFFFu));
      continue;
    }
    const double4 b = d.t_box[f0 + node];
    if (b.x - ux > 0.0 || b.y - uy > 0.0 || lx - b.z > 0.0 || ly - b.w > 0.0) continue;
    if (sp + 2 > kTreeStack) return false;
    const int2 c = d.t_child[f0 + node];
    stack[sp++] = c.y;
    stack[sp++] = c.x;
  }
  return true;
}

// One candidate (moved proxy P at move-buffer position i, partner R) of a large world's pair
// search: the rules of BroadPhase.update_pairs and find_new_contacts (exact FP64 overlap, not
// already paired, same body, active, an existing contact, Body._should_collide, the default
// filter, the contact registry), appended to the world's new-pair list.
template <typename T>
__device__ __forceinline__ void big_pair_candidate(const Dev<T>& d, int b0, int f0, const int* mpos,
                                                   const uint32_t* hash, int hcap, uint32_t* nlist, int ncap,
                                                   int* misc, int i, int P, int R, bool custom_filter) {
  if (R == P) return;
  const int mr = mpos[R];
  if (mr >= 0 && mr < i) return;  // (R, P) is already in the pair set
  const double4 fp = d.f_fat[f0 + P], fr = d.f_fat[f0 + R];
  if (!aabb_overlap(fr.x, fr.y, fr.z, fr.w, fp.x, fp.y, fp.z, fp.w)) return;
  const int4 ip = d.f_info[f0 + P], ir = d.f_info[f0 + R];
  if (ip.x == ir.x) return;  // same body
  const int tb_p = d.b_type[b0 + ip.x], tb_r = d.b_type[b0 + ir.x];
  if (!(d.b_flags[b0 + ip.x] & BF_ACTIVE) || !(d.b_flags[b0 + ir.x] & BF_ACTIVE)) return;
  {
    const uint32_t k = hkey(P, R);
    uint32_t h = hhash(k) & (uint32_t)(hcap - 1);
    for (;;) {
      const uint32_t v = hash[h];
      if (v == 0xFFFFFFFFu) break;
      if (v == k) return;
      h = (h + 1) & (uint32_t)(hcap - 1);
    }
  }
  if (tb_r != BT_DYNAMIC && tb_p != BT_DYNAMIC) return;  // Body._should_collide
  if (!custom_filter) {  // World.default_contact_filter(fixture P, fixture R)
    bool ok;
    if (ip.w == ir.w && ip.w != 0) ok = ip.w > 0;
    else ok = ((ip.z >> 16) & (ir.z & 0xFFFF)) != 0 && ((ip.z & 0xFFFF) & (ir.z >> 16)) != 0;
    if (!ok) return;
  }
  const int ta = ip.y & 255, tb = ir.y & 255;
  const bool reg = (ta == SHAPE_CIRCLE || ta == SHAPE_POLYGON || ta == SHAPE_EDGE) &&
                   (tb == SHAPE_CIRCLE || tb == SHAPE_POLYGON || tb == SHAPE_EDGE) &&
                   !(ta == SHAPE_EDGE && tb == SHAPE_EDGE);
  if (!reg) return;
  const int slot = atomicAdd(&misc[3], 1);
  if (slot < ncap) nlist[slot] = ((uint32_t)i << 16) | (uint32_t)R;
  else atomicOr(&misc[4], 1);
}

// Every moved proxy of a large world against every proxy of the world (exact
// FP64 tests), with the filters of find_new_contacts; grid (chunks, worlds).
template <typename T>
__global__ void __launch_bounds__(kBlock) k_pairs_big_find(Dev<T> d) {
  const int w = d.big_list[blockIdx.y];
  const int b0 = d.w_body_off[w];
  const int f0 = d.w_fix_off[w], nf = d.w_fix_off[w + 1] - f0;
  const int ccap = d.w_con_cap[w];
  int* misc = d.g_wmisc + 16 * w;
  const int nmove = misc[2];
  const int* mpos = d.g_mpos + f0;
  const int* mlist = d.g_mlist + f0;
  const int hcap = next_pow2(ccap + 1);
  const uint32_t* hash = d.g_hash + d.g_hbase[w];
  const int ncap = next_pow2(ccap);
  uint32_t* nlist = d.g_nlist + d.g_nbase[w];
  const long long work = (long long)nmove * nf;
  for (long long t = (long long)blockIdx.x * kBlock + threadIdx.x; t < work; t += (long long)gridDim.x * kBlock) {
    const int i = (int)(t / nf), R = (int)(t % nf);
    big_pair_candidate(d, b0, f0, mpos, hash, hcap, nlist, ncap, misc, i, mlist[i], R,
                       (d.w_flags[w] & WF_CUSTOM_FILTER) != 0);
  }
}

// The same search by traversal of the world's LBVH (built over its fat AABBs after the
// synchronisation): one lane per moved proxy visits only the proxies whose subtree boxes overlap
// its fat AABB; the candidate rules and the new-pair list are the same, and the list is put in
// canonical order afterwards, so the contacts created are identical.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_pairs_big_find_tree(Dev<T> d) {
  const int w = d.big_list[blockIdx.y];
  const int b0 = d.w_body_off[w];
  const int f0 = d.w_fix_off[w];
  const int ccap = d.w_con_cap[w];
  int* misc = d.g_wmisc + 16 * w;
  const int nmove = misc[2];
  const int* mpos = d.g_mpos + f0;
  const int* mlist = d.g_mlist + f0;
  const int hcap = next_pow2(ccap + 1);
  const uint32_t* hash = d.g_hash + d.g_hbase[w];
  const int ncap = next_pow2(ccap);
  uint32_t* nlist = d.g_nlist + d.g_nbase[w];
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < nmove; i += gridDim.x * kBlock) {
    const int P = mlist[i];
    const double4 fp = d.f_fat[f0 + P];
    const bool ok = tree_visit(d, w, fp.x, fp.y, fp.z, fp.w, [&](int R) {
      big_pair_candidate(d, b0, f0, mpos, hash, hcap, nlist, ncap, misc, i, P, R,
                         (d.w_flags[w] & WF_CUSTOM_FILTER) != 0);
    });
    if (!ok) atomicOr(&d.w_err[w], WERR_SCHEDULE);
  }
}

// Canonical order, contact creation and end-of-step bookkeeping of each large world.
template <typename T>
__global__ void __launch_bounds__(kBigBlock) k_pairs_big_finish(Dev<T> d, int finish, int do_pairs) {
  __shared__ int scan[kBigBlock];
  __shared__ int sh[4];
  const int w = d.big_list[blockIdx.x], tid = threadIdx.x;
  const int b0 = d.w_body_off[w], nb = d.w_body_off[w + 1] - b0;
  const int f0 = d.w_fix_off[w];
  const int cb = d.w_con_base[w], ccap = d.w_con_cap[w];
  int* misc = d.g_wmisc + 16 * w;
  const int* mlist = d.g_mlist + f0;
  uint32_t* nlist = d.g_nlist + d.g_nbase[w];
  int nnew = misc[3];
  if (misc[4]) {
    if (tid == 0) atomicOr(&d.w_err[w], WERR_NEWPAIRS);
    nnew = 0;
  }
  if (nnew > 0) {
    // canonical order: move-buffer position, then partner index (bitonic sort)
    const int np = next_pow2(nnew);
    for (int i = nnew + tid; i < np; i += kBigBlock) nlist[i] = 0xFFFFFFFFu;
    __syncthreads();
    for (int k = 2; k <= np; k <<= 1) {
      for (int j = k >> 1; j > 0; j >>= 1) {
        for (int i = tid; i < np; i += kBigBlock) {
          const int ixj = i ^ j;
          if (ixj > i) {
            const uint32_t a = nlist[i], b = nlist[ixj];
            const bool up = (i & k) == 0;
            if ((a > b) == up) { nlist[i] = b; nlist[ixj] = a; }
          }
        }
        __syncthreads();
      }
    }
    int used = d.w_con_used[w];
    if (used + nnew > ccap) {  // squeeze out destroyed slots
      int base = 0;
      for (int c0 = 0; c0 < used; c0 += kBigBlock) {
        const int sl = c0 + tid;
        const bool live = sl < used && (d.c_state[cb + sl] & CS_LIVE);
        int2 pr = make_int2(0, 0);
        int st = 0;
        V2<T> mat = v2<T>(0, 0);
        V4<T> lnlp{}, pts{};
        V4<double> imp{};
        uint2 id = make_uint2(0, 0);
        if (live) {
          pr = d.c_pair[cb + sl]; st = d.c_state[cb + sl]; mat = d.c_mat[cb + sl];
          lnlp = d.c_lnlp[cb + sl]; pts = d.c_pts[cb + sl]; imp = d.c_imp[cb + sl]; id = d.c_id[cb + sl];
        }
        scan[tid] = live ? 1 : 0;
        __syncthreads();
        for (int o = 1; o < kBigBlock; o <<= 1) {
          const int v = tid >= o ? scan[tid - o] : 0;
          __syncthreads();
          scan[tid] += v;
          __syncthreads();
        }
        const int dst = base + scan[tid] - (live ? 1 : 0);
        const int cnt = scan[kBigBlock - 1];
        __syncthreads();
        if (live) {
          const int s = cb + dst;
          d.c_pair[s] = pr; d.c_state[s] = st; d.c_mat[s] = mat; d.c_lnlp[s] = lnlp; d.c_pts[s] = pts;
          d.c_imp[s] = imp; d.c_id[s] = id;
        }
        base += cnt;
        __syncthreads();
      }
      used = base;
    }
    if (used + nnew > ccap) {
      if (tid == 0) atomicOr(&d.w_err[w], WERR_CAPACITY);
      nnew = 0;
    }
    for (int k = tid; k < nnew; k += kBigBlock) create_contact(d, f0, cb + used + k, mlist[nlist[k] >> 16],
                                                               (int)(nlist[k] & 0xFFFFu));
    __syncthreads();
    if (tid == 0) {
      d.w_con_used[w] = used + nnew;
      if (nnew > 0) atomicAdd(&d.stat[S_NEW], (unsigned long long)nnew);
    }
  }
  if (tid == 0) {
    if (do_pairs) d.w_move_n[w] = 0;
    d.w_new[w] = nnew;
  }
  if (!finish) return;
  if (tid == 0) {
    if (d.dt > 0.0) d.w_inv_dt0[w] = d.inv_dt;
    sh[0] = 0;
  }
  __syncthreads();
  const bool clear = !(d.w_flags[w] & WF_KEEP_FORCES);
  int asleep = 0;
  for (int i = tid; i < nb; i += kBigBlock) {
    const int gb = b0 + i;
    if (clear) {
      d.b_f[gb] = v2<T>(0, 0);
      d.b_t[gb] = (T)0;
    }
    if (d.b_type[gb] != BT_STATIC && !(d.b_flags[gb] & BF_AWAKE)) ++asleep;
  }
  if (asleep) atomicAdd(&sh[0], asleep);
  __syncthreads();
  if (tid == 0) {
    d.w_asleep[w] = sh[0];
    if (d.w_flags[w] & WF_NEW_FIXTURE) d.w_flags[w] &= ~WF_NEW_FIXTURE;
  }
}


// World.query_aabb (world.py:371-384) for query q: every fixture of world qw[q] whose fat AABB
// overlaps the box (DynamicTree.query, dynamictree.py:190-207, PyAABB.test_overlap with the
// proxy's box first), appended as (q, fixture).
template <typename T>
__global__ void __launch_bounds__(256) k_query_aabb(Dev<T> d, int nq, const int* qw, const double4* qbox, int2* out,
                                                    int cap, int* count) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nq) return;
  const int w = qw[q];
  const int f0 = d.w_fix_off[w];
  const double4 a = qbox[q];
  const bool ok = tree_visit(d, w, a.x, a.y, a.z, a.w, [&](int f) {
    const double4 b = d.f_fat[f0 + f];
    if (a.x - b.z > 0.0 || a.y - b.w > 0.0 || b.x - a.z > 0.0 || b.y - a.w > 0.0) return;
    const int slot = atomicAdd(count, 1);
    if (slot < cap) out[slot] = make_int2(q, f);
  });
  if (!ok) atomicOr(&d.w_err[w], WERR_SCHEDULE);
}

// World.ray_cast (world.py:386-440) for ray r (p1 -> p2 = seg[r]) in world rw[r]: every fixture
// whose proxy passes the tree's leaf test and whose shape the ray hits.  closest != 0: one row
// per ray, the hit of least fraction (ties: the lower fixture index), fixture -1 if none;
// otherwise every hit is appended.  Rows: (ray, fixture) and (fraction, point x, y, normal x, y),
// point = (1 - f) p1 + f p2.
template <typename T>
__global__ void __launch_bounds__(256) k_ray_cast(Dev<T> d, int nr, const int* rw, const double4* seg, int closest,
                                                  int2* out_i, double* out_d, int cap, int* count) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nr) return;
  const int w = rw[r];
  const int f0 = d.w_fix_off[w], b0 = d.w_body_off[w];
  const double4 p = seg[r];
  const RaySeg s = ray_seg(p.x, p.y, p.z, p.w);
  int best = -1;
  RayHit bh{0.0, 0.0, 0.0};
  auto emit = [&](int f, const RayHit& h) {
    const double px = (1.0 - h.fraction) * p.x + h.fraction * p.z, py = (1.0 - h.fraction) * p.y + h.fraction * p.w;
    const int slot = closest ? r : atomicAdd(count, 1);
    if (slot < cap) {
      out_i[slot] = make_int2(r, f);
      double* o = out_d + 5 * (size_t)slot;
      o[0] = h.fraction; o[1] = px; o[2] = py; o[3] = h.nx; o[4] = h.ny;
    }
  };
  const bool ok = tree_visit(d, w, s.lx, s.ly, s.ux, s.uy, [&](int f) {
    const double4 b = d.f_fat[f0 + f];
    if (!ray_leaf_passes(s, b.x, b.y, b.z, b.w)) return;
    const int4 info = d.f_info[f0 + f];
    const double4 x = d.b_xf[b0 + info.x];
    const XF64 xf{x.x, x.y, x.z, x.w};
    RayHit h;
    if (!ray_shape(geo_of(d, f0 + f), xf, p.x, p.y, p.z, p.w, h)) return;
    if (!closest) emit(f, h);
    else if (best < 0 || h.fraction < bh.fraction || (h.fraction == bh.fraction && f < best)) {
      best = f;
      bh = h;
    }
  });
  if (!ok) atomicOr(&d.w_err[w], WERR_SCHEDULE);
  if (closest) {
    if (best >= 0) emit(best, bh);
    else if (r < cap) {
      out_i[r] = make_int2(r, -1);
      double* o = out_d + 5 * (size_t)r;
      o[0] = o[1] = o[2] = o[3] = o[4] = 0.0;
    }
  }
}


// ---------------------------------------------------------------------------
// Continuous physics: World.solve_toi (world.py:594-821) after the step's final pair search,
// one lane per world in the reference's sequential order.  Each round finds the contact with
// the earliest time of impact (contacts with a bullet or a static or kinematic body; the
// lagging body's sweep advanced to the other's start first), advances its two bodies to that
// time, updates it, gathers the contacts of its dynamic bodies with static, kinematic or
// bullet bodies into a sub-step island (at most 32 contacts, 64 bodies), solves that island for
// the rest of the step (Island.solve_toi, island.py:181-243: 20 position iterations moving
// only the two bodies, then the velocity iterations without warm starting), re-synchronises
// the island bodies' proxies and runs the pair search on those moved proxies.  FP64 throughout.
// ---------------------------------------------------------------------------
constexpr int kToiContacts = 32;     // settings.py:79 (MAX_TOI_CONTACTS)
constexpr int kToiBodies = 64;       // 2 * MAX_TOI_CONTACTS
constexpr int kToiMaxSubSteps = 8;   // settings.py:73
constexpr double kToiBaumgarte = 0.75;  // settings.py:110

struct ToiPC {
  V2<double> lp0, lp1, ln, lpp, lca, lcb;
  double ra, rb, ma, mb, Ia, Ib;
  int type, npts, ia, ib;
};
struct ToiPV {
  V2<double> c, v;
  double a, w;
};
struct ToiScratch {
  VC<double> vc[kToiContacts];
  ToiPC pc[kToiContacts];
  ToiPV pv[kToiBodies];
  int bodies[kToiBodies];
  int contacts[kToiContacts];
};

template <typename T>
__device__ TSweep load_sweep(const Dev<T>& d, int gb) {
  TSweep s;
  const double2 c0 = d.b_c0[gb], c = d.b_c[gb], lc = d.b_lc[gb];
  s.c0 = v2<double>(c0.x, c0.y);
  s.c = v2<double>(c.x, c.y);
  s.lc = v2<double>(lc.x, lc.y);
  s.a0 = d.b_a0[gb];
  s.a = d.b_a[gb];
  s.alpha0 = d.b_alpha0[gb];
  return s;
}

template <typename T>
__device__ void store_sweep(const Dev<T>& d, int gb, const TSweep& s) {
  d.b_c0[gb] = make_double2(s.c0.x, s.c0.y);
  d.b_c[gb] = make_double2(s.c.x, s.c.y);
  d.b_a0[gb] = s.a0;
  d.b_a[gb] = s.a;
  d.b_alpha0[gb] = s.alpha0;
  const XF<double> xf = xf_at(s.c, s.a, s.lc);  // Body._synchronize_transform
  d.b_xf[gb] = make_double4(xf.p.x, xf.p.y, xf.q.c, xf.q.s);
}

template <typename T>
__device__ void body_advance(const Dev<T>& d, int gb, double alpha) {  // Body._advance (body.py:930-937)
  TSweep s = load_sweep(d, gb);
  sweep_advance(s, alpha);
  s.c = s.c0;
  s.a = s.a0;
  store_sweep(d, gb, s);
}

// ContactSolver.solve_toi_position_constraints (contact.py:866-933): only the TOI bodies
// (island indices 0 and 1) move.
__device__ inline bool toi_position_sweep(ToiScratch& S, int nc) {
  double min_sep = 0.0;
  for (int k = 0; k < nc; ++k) {
    const ToiPC& pc = S.pc[k];
    V2<double> ca = S.pv[pc.ia].c, cb = S.pv[pc.ib].c;
    double aa = S.pv[pc.ia].a, ab = S.pv[pc.ib].a;
    const bool mva = pc.ia < 2, mvb = pc.ib < 2;
    const double ima = mva ? pc.ma : 0.0, iIa = mva ? pc.Ia : 0.0;
    const double imb = mvb ? pc.mb : 0.0, iIb = mvb ? pc.Ib : 0.0;
    for (int j = 0; j < pc.npts; ++j) {
      const XF<double> xfa = xf_at(ca, aa, pc.lca), xfb = xf_at(cb, ab, pc.lcb);
      V2<double> normal, point;
      double sep;
      const V2<double> lpj = j == 0 ? pc.lp0 : pc.lp1;
      if (pc.type == M_CIRCLES) {  // PositionSolverManifold (contact_util.py)
        const V2<double> pa = xmul(xfa, pc.lpp), pb = xmul(xfb, pc.lp0);
        normal = normalized(pb - pa);
        point = scale(0.5, pa + pb);
        sep = dot(pb - pa, normal) - pc.ra - pc.rb;
      } else if (pc.type == M_FACE_A) {
        normal = rmul(xfa.q, pc.ln);
        const V2<double> plane = xmul(xfa, pc.lpp), clip = xmul(xfb, lpj);
        sep = dot(clip - plane, normal) - pc.ra - pc.rb;
        point = clip;
      } else {
        normal = rmul(xfb.q, pc.ln);
        const V2<double> plane = xmul(xfb, pc.lpp), clip = xmul(xfa, lpj);
        sep = dot(clip - plane, normal) - pc.ra - pc.rb;
        point = clip;
        normal = -normal;
      }
      const V2<double> ra = point - ca, rb = point - cb;
      min_sep = pmin(min_sep, sep);
      const double C = pclamp(kToiBaumgarte * (sep + LINEAR_SLOP), -MAX_LINEAR_CORRECTION, 0.0);
      const double rna = sq(cross(ra, normal)), rnb = sq(cross(rb, normal));
      const double K = ima + imb + iIa * rna + iIb * rnb;
      const double impulse = K > 0.0 ? -C / K : 0.0;
      const V2<double> P = scale(impulse, normal);
      ca = ca - scale(ima, P);
      aa -= iIa * cross(ra, P);
      cb = cb + scale(imb, P);
      ab += iIb * cross(rb, P);
    }
    S.pv[pc.ia].c = ca;
    S.pv[pc.ia].a = aa;
    S.pv[pc.ib].c = cb;
    S.pv[pc.ib].a = ab;
  }
  return min_sep >= -1.5 * LINEAR_SLOP;
}

// ContactSolver.initialize_velocity_constraints (contact.py:414-496) of island contact k at
// the island's positions, in world coordinates.
__device__ inline void toi_init_velocity(ToiScratch& S, int k) {
  VC<double>& vc = S.vc[k];
  const ToiPC& pc = S.pc[k];
  const ToiPV &A = S.pv[pc.ia], &B = S.pv[pc.ib];
  const XF<double> xfa = xf_at(A.c, A.a, pc.lca), xfb = xf_at(B.c, B.a, pc.lcb);
  V2<double> w0 = v2<double>(0, 0), w1 = v2<double>(0, 0);
  const V2<double> normal = world_manifold<double>(pc.type, pc.npts, pc.ln, pc.lpp, pc.lp0, pc.lp1, xfa, pc.ra, xfb,
                                                   pc.rb, w0, w1);
  vc.nx = normal.x;
  vc.ny = normal.y;
  const double ma = vc.ma, mb = vc.mb, Ia = vc.Ia, Ib = vc.Ib;
  for (int j = 0; j < vc.npts; ++j) {
    VCP<double>& p = vc.p[j];
    const V2<double> wp = j == 0 ? w0 : w1;
    const V2<double> ra = wp - A.c, rb = wp - B.c;
    p.rax = ra.x; p.ray = ra.y; p.rbx = rb.x; p.rby = rb.y;
    const double kn = ma + mb + Ia * sq(cross(ra, normal)) + Ib * sq(cross(rb, normal));
    p.nm = kn > 0.0 ? 1.0 / kn : 0.0;
    const V2<double> tangent = cross_vs(normal, 1.0);
    const double kt = ma + mb + Ia * sq(cross(ra, tangent)) + Ib * sq(cross(rb, tangent));
    p.tm = kt > 0.0 ? 1.0 / kt : 0.0;
    p.bias = 0.0;
    const double vrel = dot(normal, B.v + cross_sv(B.w, rb) - A.v - cross_sv(A.w, ra));
    if (vrel < -VELOCITY_THRESHOLD) p.bias = -vc.restitution * vrel;
  }
  vc.k11 = 1; vc.k12 = 0; vc.k22 = 1; vc.n11 = 1; vc.n12 = 0; vc.n21 = 0; vc.n22 = 1;
  if (vc.npts == 2) {
    const V2<double> ra1 = dv2(vc.p[0].rax, vc.p[0].ray), rb1 = dv2(vc.p[0].rbx, vc.p[0].rby);
    const V2<double> ra2 = dv2(vc.p[1].rax, vc.p[1].ray), rb2 = dv2(vc.p[1].rbx, vc.p[1].rby);
    const double rn1a = cross(ra1, normal), rn1b = cross(rb1, normal);
    const double rn2a = cross(ra2, normal), rn2b = cross(rb2, normal);
    const double k11 = ma + mb + Ia * rn1a * rn1a + Ib * rn1b * rn1b;
    const double k22 = ma + mb + Ia * rn2a * rn2a + Ib * rn2b * rn2b;
    const double k12 = ma + mb + Ia * rn1a * rn2a + Ib * rn1b * rn2b;
    if (sq(k11) < 1000.0 * (k11 * k22 - sq(k12))) {
      vc.k11 = k11; vc.k12 = k12; vc.k22 = k22;
      double det = k11 * k22 - k12 * k12;
      if (det != 0.0) det = 1.0 / det;
      vc.n11 = det * k22; vc.n21 = -det * k12; vc.n12 = -det * k12; vc.n22 = det * k11;
    } else {
      vc.npts = 1;  // quirk Q6
    }
  }
}

// One World.solve_toi of world w (lane 0 of its wavefront).
template <typename T>
__device__ void world_solve_toi(const Dev<T>& d, int w, ToiScratch& S) {
  const int b0 = d.w_body_off[w], nb = d.w_body_off[w + 1] - b0;
  const int f0 = d.w_fix_off[w], nf = d.w_fix_off[w + 1] - f0;
  const int cb = d.w_con_base[w], ccap = d.w_con_cap[w];
  int* mv = d.toi_move + f0;
  int seq = 0;  // listener records of this phase, in call order
  unsigned long long calls = 0, events = 0;
  for (int i = 0; i < nb; ++i) {
    d.b_alpha0[b0 + i] = 0.0;
    d.b_tflag[b0 + i] = 0;
  }
  for (int sl = 0; sl < d.w_con_used[w]; ++sl) {
    d.c_toi[cb + sl] = 1.0;
    d.c_tmeta[cb + sl] = 0;
  }
  const double eps = NormEps<double>::value;
  for (;;) {
    const int used = d.w_con_used[w];
    int min_s = -1;
    double min_alpha = 1.0;
    for (int sl = 0; sl < used; ++sl) {
      const int s = cb + sl, st = d.c_state[s];
      if (!(st & CS_LIVE) || !(st & CS_ENABLED)) continue;
      const int meta = d.c_tmeta[s];
      if ((meta >> 8) > kToiMaxSubSteps) continue;
      double alpha = 1.0;
      if (meta & 1) {
        alpha = d.c_toi[s];
      } else {
        const int2 pr = d.c_pair[s];
        const int fa = f0 + pr.x, fb = f0 + pr.y;
        const int ba = b0 + d.f_info[fa].x, bb = b0 + d.f_info[fb].x;
        const int ta = d.b_type[ba], tb = d.b_type[bb], fla = d.b_flags[ba], flb = d.b_flags[bb];
        const bool active_a = (fla & BF_AWAKE) && ta != BT_STATIC, active_b = (flb & BF_AWAKE) && tb != BT_STATIC;
        if (!active_a && !active_b) continue;
        const bool col_a = (fla & BF_BULLET) || ta != BT_DYNAMIC, col_b = (flb & BF_BULLET) || tb != BT_DYNAMIC;
        if (!col_a && !col_b) continue;
        double alpha0 = d.b_alpha0[ba];
        if (d.b_alpha0[ba] < d.b_alpha0[bb]) {
          alpha0 = d.b_alpha0[bb];
          TSweep sw = load_sweep(d, ba);
          sweep_advance(sw, alpha0);
          store_sweep(d, ba, sw);
        } else if (d.b_alpha0[bb] < d.b_alpha0[ba]) {
          TSweep sw = load_sweep(d, bb);
          sweep_advance(sw, alpha0);
          store_sweep(d, bb, sw);
        }
        TProxy pa, pb;
        tproxy_init(pa, geo_of(d, fa));
        tproxy_init(pb, geo_of(d, fb));
        double beta;
        int state = time_of_impact(pa, pb, load_sweep(d, ba), load_sweep(d, bb), beta);
        ++calls;
        if (state == TOI_STUCK) atomicOr(&d.w_err[w], WERR_TOI);
        alpha = state == TOI_TOUCHING ? pmin(alpha0 + (1.0 - alpha0) * beta, 1.0) : 1.0;
        d.c_toi[s] = alpha;
        d.c_tmeta[s] = meta | 1;
      }
      if (alpha < min_alpha) {
        min_s = s;
        min_alpha = alpha;
      }
    }
    if (min_s < 0 || (1.0 - 10.0 * eps) < min_alpha) break;
    // advance the two bodies to the time of impact and update their contact
    const int2 mpr = d.c_pair[min_s];
    const int ba = b0 + d.f_info[f0 + mpr.x].x, bb = b0 + d.f_info[f0 + mpr.y].x;
    const TSweep back_a = load_sweep(d, ba), back_b = load_sweep(d, bb);
    body_advance(d, ba, min_alpha);
    body_advance(d, bb, min_alpha);
    int st = update_contact(d, w, min_s, d.c_state[min_s], &seq);
    d.c_state[min_s] = st;
    const int m0 = d.c_tmeta[min_s];  // its cached time is spent, one more sub-step
    d.c_tmeta[min_s] = (m0 & 2) | (((m0 >> 8) + 1) << 8);
    if (!(st & CS_ENABLED) || !(st & CS_TOUCHING)) {
      d.c_state[min_s] = st & ~CS_ENABLED;
      store_sweep(d, ba, back_a);
      store_sweep(d, bb, back_b);
      continue;
    }
    wake_body(d, ba);
    wake_body(d, bb);
    // the sub-step island: the two bodies, their contact, and the contacts of their dynamic
    // bodies with static, kinematic or bullet bodies (in each body's contact order)
    int nbi = 0, nci = 0;
    S.bodies[nbi++] = ba;
    d.b_tflag[ba] = nbi;
    S.bodies[nbi++] = bb;
    d.b_tflag[bb] = nbi;
    S.contacts[nci++] = min_s;
    d.c_tmeta[min_s] |= 2;
    for (int side = 0; side < 2; ++side) {
      const int body = side == 0 ? ba : bb;
      if (d.b_type[body] != BT_DYNAMIC) continue;
      const bool bullet_b = (d.b_flags[body] & BF_BULLET) != 0;
      for (int sl = 0; sl < d.w_con_used[w]; ++sl) {
        const int s = cb + sl;
        if (!(d.c_state[s] & CS_LIVE)) continue;
        const int2 pr = d.c_pair[s];
        const int ga = b0 + d.f_info[f0 + pr.x].x, gbd = b0 + d.f_info[f0 + pr.y].x;
        if (ga != body && gbd != body) continue;
        if (nbi == kToiBodies || nci == kToiContacts) break;  // the island's capacity
        if (d.c_tmeta[s] & 2) continue;
        const int other = ga == body ? gbd : ga;
        if (d.b_type[other] == BT_DYNAMIC && !bullet_b && !(d.b_flags[other] & BF_BULLET)) continue;
        const TSweep back = load_sweep(d, other);
        if (!d.b_tflag[other]) body_advance(d, other, min_alpha);
        const int cst = update_contact(d, w, s, d.c_state[s], &seq);
        d.c_state[s] = cst;
        if (!(cst & CS_ENABLED) || !(cst & CS_TOUCHING)) {
          store_sweep(d, other, back);
          continue;
        }
        d.c_tmeta[s] |= 2;
        S.contacts[nci++] = s;
        if (d.b_tflag[other]) continue;
        if (d.b_type[other] != BT_STATIC) wake_body(d, other);
        S.bodies[nbi++] = other;
        d.b_tflag[other] = nbi;
      }
    }
    // Island.solve_toi over the rest of the step
    const double sub_dt = (1.0 - min_alpha) * d.dt;
    for (int i = 0; i < nbi; ++i) {
      const int gb = S.bodies[i];
      const double2 c = d.b_c[gb];
      S.pv[i].c = v2<double>(c.x, c.y);
      S.pv[i].a = d.b_a[gb];
      S.pv[i].v = d.b_v[gb];
      S.pv[i].w = d.b_w[gb];
    }
    for (int k = 0; k < nci; ++k) {  // ContactSolver.__init__ (contact.py:376-412), no warm starting
      const int s = S.contacts[k], cst = d.c_state[s];
      const int2 pr = d.c_pair[s];
      const int fa = f0 + pr.x, fb = f0 + pr.y;
      const int ga = b0 + d.f_info[fa].x, gbd = b0 + d.f_info[fb].x;
      const V2<T> mat = d.c_mat[s];
      const V4<T> lnlp = d.c_lnlp[s], pts = d.c_pts[s];
      const V4<double> imp = d.c_imp[s];
      VC<double>& vc = S.vc[k];
      ToiPC& pc = S.pc[k];
      vc.friction = (double)mat.x;
      vc.restitution = (double)mat.y;
      vc.ia = d.b_tflag[ga] - 1;
      vc.ib = d.b_tflag[gbd] - 1;
      vc.ma = (double)d.b_im[ga];
      vc.mb = (double)d.b_im[gbd];
      vc.Ia = (double)d.b_iI[ga];
      vc.Ib = (double)d.b_iI[gbd];
      vc.npts = (cst >> CS_PC_SHIFT) & 3;
      vc.rka = vc.rkb = 0;
      vc.nxt = 0u;
      vc.p[0].ni = imp.x; vc.p[0].ti = imp.y;  // the stored impulses (quirk Q8)
      vc.p[1].ni = imp.z; vc.p[1].ti = imp.w;
      vc.p[0].pad = vc.p[1].pad = 0.0;
      pc.lp0 = dv2(pts.x, pts.y);
      pc.lp1 = dv2(pts.z, pts.w);
      pc.ln = dv2(lnlp.x, lnlp.y);
      pc.lpp = dv2(lnlp.z, lnlp.w);
      const double2 lca = d.b_lc[ga], lcb = d.b_lc[gbd];
      pc.lca = v2<double>(lca.x, lca.y);
      pc.lcb = v2<double>(lcb.x, lcb.y);
      pc.ra = d.f_rad64[fa];
      pc.rb = d.f_rad64[fb];
      pc.ma = vc.ma; pc.mb = vc.mb; pc.Ia = vc.Ia; pc.Ib = vc.Ib;
      pc.type = (cst >> CS_MTYPE_SHIFT) & 3;
      pc.npts = vc.npts;
      pc.ia = vc.ia;
      pc.ib = vc.ib;
    }
    for (int it = 0; it < 20; ++it)
      if (toi_position_sweep(S, nci)) break;
    {  // the two TOI bodies' sweeps start at the solved positions ("leap of faith")
      TSweep sa = load_sweep(d, ba), sb = load_sweep(d, bb);
      sa.c0 = S.pv[0].c; sa.a0 = S.pv[0].a;
      sb.c0 = S.pv[1].c; sb.a0 = S.pv[1].a;
      d.b_c0[ba] = make_double2(sa.c0.x, sa.c0.y);
      d.b_a0[ba] = sa.a0;
      d.b_c0[bb] = make_double2(sb.c0.x, sb.c0.y);
      d.b_a0[bb] = sb.a0;
    }
    for (int k = 0; k < nci; ++k) toi_init_velocity(S, k);
    for (int it = 0; it < d.vel_iters; ++it)
      for (int k = 0; k < nci; ++k) {
        VC<double>& vc = S.vc[k];
        V2<double> va = S.pv[vc.ia].v, vb = S.pv[vc.ib].v;
        double wa = S.pv[vc.ia].w, wbv = S.pv[vc.ib].w;
        solve_contact_velocity(vc, va, wa, vb, wbv);
        S.pv[vc.ia].v = va; S.pv[vc.ia].w = wa;
        S.pv[vc.ib].v = vb; S.pv[vc.ib].w = wbv;
      }
    for (int i = 0; i < nbi; ++i) {  // integrate positions
      const int gb = S.bodies[i];
      V2<double> v = S.pv[i].v;
      double wv = S.pv[i].w;
      const V2<double> tr = scale(sub_dt, v);
      if (tr.x * tr.x + tr.y * tr.y > MAX_TRANSLATION * MAX_TRANSLATION)
        v = scale(MAX_TRANSLATION / sqrt(tr.x * tr.x + tr.y * tr.y), v);
      const double rotation = sub_dt * wv;
      if (rotation * rotation > MAX_ROTATION * MAX_ROTATION) wv *= MAX_ROTATION / fabs(rotation);
      const V2<double> c = S.pv[i].c + scale(sub_dt, v);
      const double a = S.pv[i].a + sub_dt * wv;
      TSweep sw = load_sweep(d, gb);
      sw.c = c;
      sw.a = a;
      store_sweep(d, gb, sw);
      d.b_v[gb] = v;
      d.b_w[gb] = wv;
    }
    if (d.ev_mask & (1 << EV_POST_SOLVE))  // Island.report
      for (int k = 0; k < nci; ++k) {
        const int e = ev_open(d, EV_POST_SOLVE);
        if (e < 0) continue;
        const VC<double>& vc = S.vc[k];
        ev_write(d, e, w, EV_POST_SOLVE, 3, seq++, S.contacts[k], d.c_state[S.contacts[k]]);
        int* ei = d.ev_i + (size_t)e * B2G_EV_NI;
        double* ed = d.ev_d + (size_t)e * B2G_EV_ND;
        ei[B2G_EI_OLD_PC] = vc.npts;
        ed[B2G_ED_OLD + B2G_ED_P0_NI] = vc.p[0].ni;
        ed[B2G_ED_OLD + B2G_ED_P0_TI] = vc.p[0].ti;
        ed[B2G_ED_OLD + B2G_ED_P1_NI] = vc.npts >= 2 ? vc.p[1].ni : 0.0;
        ed[B2G_ED_OLD + B2G_ED_P1_TI] = vc.npts >= 2 ? vc.p[1].ti : 0.0;
      }
    ++events;
    // re-synchronise the island's dynamic bodies, forget their contacts' times of impact
    int nmove = 0;
    for (int i = 0; i < nbi; ++i) {
      const int gb = S.bodies[i];
      d.b_tflag[gb] = 0;
      if (d.b_type[gb] != BT_DYNAMIC) continue;
      const int lb = gb - b0;
      for (int f = 0; f < nf; ++f)
        if (d.f_info[f0 + f].x == lb && sync_fixture_aabb(d, f0 + f, gb)) mv[nmove++] = f;
      for (int sl = 0; sl < d.w_con_used[w]; ++sl) {
        const int s = cb + sl;
        if (!(d.c_state[s] & CS_LIVE)) continue;
        const int2 pr = d.c_pair[s];
        if (b0 + d.f_info[f0 + pr.x].x == gb || b0 + d.f_info[f0 + pr.y].x == gb) d.c_tmeta[s] &= ~3;
      }
    }
    // ContactManager.find_new_contacts over the moved proxies (move order, then partner index)
    for (int i = 0; i < nmove; ++i) {
      const int P = mv[i];
      const double4 fp = d.f_fat[f0 + P];
      const int4 ip = d.f_info[f0 + P];
      for (int R = 0; R < nf; ++R) {
        if (R == P) continue;
        bool earlier = false;
        for (int j = 0; j < i; ++j) earlier |= mv[j] == R;
        if (earlier) continue;  // (R, P) is in the pair set already
        const double4 fr = d.f_fat[f0 + R];
        if (!aabb_overlap(fr.x, fr.y, fr.z, fr.w, fp.x, fp.y, fp.z, fp.w)) continue;
        const int4 ir = d.f_info[f0 + R];
        if (ip.x == ir.x) continue;
        const int gp = b0 + ip.x, gr = b0 + ir.x;
        if (!(d.b_flags[gp] & BF_ACTIVE) || !(d.b_flags[gr] & BF_ACTIVE)) continue;
        bool exists = false;
        const int used2 = d.w_con_used[w];
        for (int sl = 0; sl < used2 && !exists; ++sl) {
          const int s = cb + sl;
          if (!(d.c_state[s] & CS_LIVE)) continue;
          const int2 pr = d.c_pair[s];
          exists = (pr.x == P && pr.y == R) || (pr.x == R && pr.y == P);
        }
        if (exists) continue;
        if (d.b_type[gp] != BT_DYNAMIC && d.b_type[gr] != BT_DYNAMIC) continue;  // Body._should_collide
        if (joint_blocks(d, w, ip.x, ir.x)) continue;
        bool ok;
        if (ip.w == ir.w && ip.w != 0) ok = ip.w > 0;  // World.default_contact_filter
        else ok = ((ip.z >> 16) & (ir.z & 0xFFFF)) != 0 && ((ip.z & 0xFFFF) & (ir.z >> 16)) != 0;
        if (!ok) continue;
        const int ta = ip.y & 255, tb = ir.y & 255;
        if (ta == SHAPE_EDGE && tb == SHAPE_EDGE) continue;  // no collider registered
        if (used2 >= ccap) {
          atomicOr(&d.w_err[w], WERR_CAPACITY);
          continue;
        }
        create_contact(d, f0, cb + used2, P, R);
        d.c_toi[cb + used2] = 0.0;
        d.c_tmeta[cb + used2] = 0;
        d.w_con_used[w] = used2 + 1;
      }
    }
  }
  if (calls) atomicAdd(&d.stat[S_TOI_CALLS], calls);
  if (events) {
    atomicAdd(&d.stat[S_TOI_EVENTS], events);
    int asleep = 0;  // the events woke bodies
    for (int i = 0; i < nb; ++i)
      if (d.b_type[b0 + i] != BT_STATIC && !(d.b_flags[b0 + i] & BF_AWAKE)) ++asleep;
    d.w_asleep[w] = asleep;
  }
}

template <typename T>
__global__ void __launch_bounds__(64) k_toi(Dev<T> d, int nworlds) {
  const int k = blockIdx.x;
  if (k >= nworlds || threadIdx.x != 0) return;
  const int w = d.toi_worlds[k];
  if (!(d.w_flags[w] & WF_CONTINUOUS) || !(d.dt > 0.0)) return;
  world_solve_toi(d, w, *(ToiScratch*)(d.toi_scr + (size_t)k * sizeof(ToiScratch)));
}

template <typename T>
__global__ void k_reset_queue(Dev<T> d) {
  if (blockIdx.x == 0) {
    if (threadIdx.x == 0) *d.q_count = 0;
    if (threadIdx.x == 0 && d.ev_count) *d.ev_count = 0;  // events are those of the last step
    if (threadIdx.x < kQWaveBase) d.q_cls[threadIdx.x] = 0;
  }
}

// b2g_pair_search ran the new-fixture search ahead of the step: the flag is done with
// (World.step clears _new_fixture right after that search, world.py:316-319).
template <typename T>
__global__ void k_clear_new_fixture(Dev<T> d) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w < d.nw) d.w_flags[w] &= ~WF_NEW_FIXTURE;
}

// b2g_new_contacts: row r of the listing is new contact r - off[w] of world w, where
// off[w] <= r < off[w + 1] (off: exclusive prefix of w_new, nw + 1 entries).  The new
// contacts of a world are the last w_new[w] slots of its list.
template <typename T>
__global__ void k_new_contacts(Dev<T> d, const int* off, int total, int* out) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= total) return;
  int lo = 0, hi = d.nw;  // off[lo] <= r < off[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= r) lo = mid;
    else hi = mid;
  }
  const int w = lo, j = r - off[w];
  const int s = d.w_con_base[w] + d.w_con_used[w] - d.w_new[w] + j;
  const int2 pr = d.c_pair[s];
  const int st = d.c_state[s];
  out[4 * r] = w;
  out[4 * r + 1] = pr.x;
  out[4 * r + 2] = pr.y;
  out[4 * r + 3] = ((st & CS_LIVE) ? 0 : B2G_NC_DROPPED) | ((st & CS_SWAPPED) ? (int)B2G_CF_SWAPPED : 0);
}

// b2g_drop_contacts: (world, j) pairs, j < w_new[world] (checked by the host)
template <typename T>
__global__ void k_drop_new(Dev<T> d, const int* wj, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int w = wj[2 * i], j = wj[2 * i + 1];
  d.c_state[d.w_con_base[w] + d.w_con_used[w] - d.w_new[w] + j] = 0;
}

// b2g_disable_contacts: contact.enabled = False for (world, slot) pairs (slot < w_con_used[world],
// checked by the host); a destroyed slot stays as it is
template <typename T>
__global__ void k_disable(Dev<T> d, const int* ws, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int s = d.w_con_base[ws[2 * i]] + ws[2 * i + 1];
  const int st = d.c_state[s];
  if (st & CS_LIVE) d.c_state[s] = st & ~CS_ENABLED;
}

// Orders the island queue by size class, largest first, so that the lanes of
// a wave of k_solve carry islands of similar length.  Each block takes a
// contiguous chunk of the queue, counts its classes in LDS and reserves its
// ranges with one global atomic per class.
template <typename T>
__global__ void k_order_islands(Dev<T> d) {
  __shared__ int base[kSizeClasses], lcnt[kSizeClasses], lbase[kSizeClasses];
  if (threadIdx.x < kSizeClasses) lcnt[threadIdx.x] = 0;
  if (threadIdx.x == 0) {
    int run = 0, waves = 0;
    for (int c = kSizeClasses - 1; c >= 0; --c) {
      base[c] = run;
      const int nc = d.q_cls[c];
      if (blockIdx.x == 0) {
        d.q_cls[kQBase + c] = run;
        d.q_cls[kQWaveBase + c] = waves;
        if (c >= d.heavy_class)
          waves += (nc * team_size(c, d.team_full_class, d.team_max) + kSolveBlock - 1) / kSolveBlock;
      }
      run += nc;
      if (c == d.heavy_class && blockIdx.x == 0) {
        d.q_cls[kQHeavy] = run;
        d.q_cls[kQHeavyWaves] = waves;
      }
    }
  }
  __syncthreads();
  const int n = *d.q_count;
  const int per = (n + gridDim.x - 1) / gridDim.x;
  const int q0 = blockIdx.x * per, q1 = min(n, q0 + per);
  for (int q = q0 + threadIdx.x; q < q1; q += blockDim.x)
    atomicAdd(&lcnt[queue_class(d, d.q_nc[q] & 0x3FFFFFFF, d.q_isl[q].z & 0xFFFF)], 1);
  __syncthreads();
  if (threadIdx.x < kSizeClasses) {
    const int c = threadIdx.x;
    lbase[c] = lcnt[c] > 0 ? base[c] + atomicAdd(&d.q_cls[kSizeClasses + c], lcnt[c]) : 0;
    lcnt[c] = 0;
  }
  __syncthreads();
  for (int q = q0 + threadIdx.x; q < q1; q += blockDim.x) {
    const int nc = d.q_nc[q];
    const int4 isl = d.q_isl[q];
    const int c = queue_class(d, nc & 0x3FFFFFFF, isl.z & 0xFFFF);
    const int pos = lbase[c] + atomicAdd(&lcnt[c], 1);
    d.qs_isl[pos] = isl;
    d.qs_nc[pos] = nc;
  }
}

// Orders each team class's range of the queue by the passes its islands need,
// longest first: a team lane walks its sweeps in steps of the island's sweep
// period, and owns a team size's share of the contacts of each sweep, so the
// larger of the two estimates the island's length.  When a class holds more
// islands than the chip has wavefront slots, its last wavefronts to start
// then carry its shortest islands.  One workgroup per class, a counting sort
// through the unordered queue (free once k_order_islands has read it).
// Scheduling only: results do not depend on it.
constexpr int kTeamSortBins = 2048;
template <typename T>
__global__ void __launch_bounds__(kPrefixBlock) k_order_team_islands(Dev<T> d) {
  __shared__ int bins[kTeamSortBins], scan[kPrefixBlock];
  const int c = blockIdx.x;
  if (c < d.heavy_class) return;
  const int b = d.q_cls[kQBase + c], n = d.q_cls[c];
  if (n < 2) return;
  const int TS = team_size(c, d.team_full_class, d.team_max);
  auto bin_of = [&](const int4& in, int nc) {
    const int est = team_island_length(in, nc, TS);
    return kTeamSortBins - 1 - (est < kTeamSortBins - 1 ? est : kTeamSortBins - 1);
  };
  for (int i = threadIdx.x; i < kTeamSortBins; i += kPrefixBlock) bins[i] = 0;
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += kPrefixBlock) atomicAdd(&bins[bin_of(d.qs_isl[b + i], d.qs_nc[b + i])], 1);
  __syncthreads();
  block_exclusive_scan<kPrefixBlock>(bins, kTeamSortBins, scan);
  for (int i = threadIdx.x; i < n; i += kPrefixBlock) {
    const int4 in = d.qs_isl[b + i];
    const int nc = d.qs_nc[b + i];
    const int pos = b + atomicAdd(&bins[bin_of(in, nc)], 1);
    d.q_isl[pos] = in;
    d.q_nc[pos] = nc;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += kPrefixBlock) {
    d.qs_isl[b + i] = d.q_isl[b + i];
    d.qs_nc[b + i] = d.q_nc[b + i];
  }
}

// ---------------------------------------------------------------------------
// parity seams
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_narrow_seam(int n, const double* geo_a, const int* info_a, const double* geo_b, const int* info_b,
                              const double* xfa, const double* xfb, int* out_i, T* out_t) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  Geo<double, double> A, B;  // FP64 geometry and evaluation, as the step does; the manifold is stored as T
  A.g = geo_a + (size_t)k * GEO_STRIDE;
  A.type = info_a[k] & 255;
  A.n = (info_a[k] >> 8) & 255;
  A.has0 = (info_a[k] >> 16) & 1;
  A.has3 = (info_a[k] >> 17) & 1;
  B.g = geo_b + (size_t)k * GEO_STRIDE;
  B.type = info_b[k] & 255;
  B.n = (info_b[k] >> 8) & 255;
  B.has0 = (info_b[k] >> 16) & 1;
  B.has3 = (info_b[k] >> 17) & 1;
  XF<double> ta, tb;
  ta.p = v2<double>(0, 0);
  ta.q.c = xfa[4 * k + 2];
  ta.q.s = xfa[4 * k + 3];
  tb.p = v2<double>(xfb[4 * k] - xfa[4 * k], xfb[4 * k + 1] - xfa[4 * k + 1]);
  tb.q.c = xfb[4 * k + 2];
  tb.q.s = xfb[4 * k + 3];
  Man<double> m;
  m.type = 0; m.pc = 0;
  m.ln = m.lp = m.p0 = m.p1 = v2<double>(0, 0);
  m.id0 = m.id1 = 0;
  bool ok = evaluate(m, A, ta, B, tb);
  out_i[4 * k] = ok ? m.type : -1;
  out_i[4 * k + 1] = m.pc;
  out_i[4 * k + 2] = (int)m.id0;
  out_i[4 * k + 3] = (int)m.id1;
  T* o = out_t + (size_t)k * 8;
  o[0] = (T)m.ln.x; o[1] = (T)m.ln.y; o[2] = (T)m.lp.x; o[3] = (T)m.lp.y;
  o[4] = (T)m.p0.x; o[5] = (T)m.p0.y; o[6] = (T)m.p1.x; o[7] = (T)m.p1.y;
}

// pair seam: exact FP64 fat-AABB tests of a move buffer against n proxies;
// one lane per (move entry, proxy) with the pair-set dedup rule.
__global__ void k_pairs_seam(int n, const double4* fat, int nmove, const int* move, const int* mpos, uint32_t* out,
                             int cap, int* count) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)nmove * n) return;
  const int i = (int)(t / n), R = (int)(t % n);
  const int P = move[i];
  if (R == P) return;
  const int mr = mpos[R];
  if (mr >= 0 && mr < i) return;
  double4 fp = fat[P], fr = fat[R];
  if (!aabb_overlap(fr.x, fr.y, fr.z, fr.w, fp.x, fp.y, fp.z, fp.w)) return;
  int slot = atomicAdd(count, 1);
  if (slot < cap) {
    out[2 * slot] = (uint32_t)i;
    out[2 * slot + 1] = (uint32_t)R;
  }
}


// ---------------------------------------------------------------------------
// The pair seam without allocation (b2g_pairs_noalloc): device inputs and outputs, its
// scratch (move positions, per-moved-proxy pair counts and output offsets, an error word and
// the total) in the file-scope array below.  The host refers to the array only through the
// address hipGetSymbolAddress returns for the current device.
// ---------------------------------------------------------------------------
constexpr int kPdMaxProxies = 1 << 20;  // largest n
constexpr int kPdMaxMoved = 1 << 20;    // largest move buffer
constexpr int kPdMisc = 64;             // [0] error bits (1: move entry out of range, 2: duplicate)
constexpr size_t kPdCounts = (size_t)kPdMaxProxies, kPdOffsets = kPdCounts + kPdMaxMoved,
                 kPdMiscAt = kPdOffsets + kPdMaxMoved, kPdScratchInts = kPdMiscAt + kPdMisc;
static __device__ __align__(256) int scratch[kPdScratchInts];  // 8 MiB + 256 B per device

__global__ void k_pd_init(int n, int* sc) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) sc[i] = -1;
  if (blockIdx.x == 0 && threadIdx.x < kPdMisc) sc[kPdMiscAt + threadIdx.x] = 0;
}

__global__ void k_pd_mark(int n, int nmove, const int* move, int* sc) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nmove; k += gridDim.x * blockDim.x) {
    const int P = move[k];
    if (P < 0 || P >= n) {
      atomicOr(&sc[kPdMiscAt], 1);
      continue;
    }
    if (atomicCAS(&sc[P], -1, k) != -1) atomicOr(&sc[kPdMiscAt], 2);
  }
}

// One workgroup per moved proxy i (move-buffer position): its partners R in ascending order,
// under the same rule as k_pairs_seam (R != P, R not earlier in the move buffer, inclusive
// FP64 overlap).  kWrite = false counts them; kWrite = true writes (P, R) from the proxy's
// offset on, so the output is in canonical order without a sort.
template <bool kWrite>
__global__ void __launch_bounds__(256) k_pd_pairs(int n, const double4* fat, int nmove, const int* move,
                                                  const int* sc_ro, int* sc, int* out, int cap) {
  __shared__ int wsum[4];
  const int i = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  if (i >= nmove || sc_ro[kPdMiscAt] != 0) return;  // the whole workgroup
  const int P = move[i];
  const double4 fp = fat[P];
  const int base = kWrite ? sc_ro[kPdOffsets + i] : 0;
  int run = 0;
  for (int r0 = 0; r0 < n; r0 += 256) {
    const int R = r0 + tid;
    bool hit = false;
    if (R < n && R != P) {
      const int mr = sc_ro[R];
      if (!(mr >= 0 && mr < i)) {
        const double4 fr = fat[R];
        hit = aabb_overlap(fr.x, fr.y, fr.z, fr.w, fp.x, fp.y, fp.z, fp.w);
      }
    }
    const unsigned long long b = __ballot(hit);
    if (lane == 0) wsum[wv] = __popcll(b);
    __syncthreads();
    int before = run + __popcll(b & ((1ull << lane) - 1));
    for (int j = 0; j < wv; ++j) before += wsum[j];
    if (kWrite && hit && base + before < cap) {
      out[2 * (base + before)] = P;
      out[2 * (base + before) + 1] = R;
    }
    run += wsum[0] + wsum[1] + wsum[2] + wsum[3];
    __syncthreads();
  }
  if (!kWrite && tid == 0) sc[kPdCounts + i] = run;
}

// Exclusive scan of the per-proxy counts into offsets (one workgroup); *count = the total, or
// -1 when the move buffer was invalid.
__global__ void __launch_bounds__(256) k_pd_scan(int nmove, int* sc, int* count) {
  __shared__ int part[256];
  const int tid = threadIdx.x;
  const int per = (nmove + 255) / 256, lo = min(nmove, tid * per), hi = min(nmove, lo + per);
  int sum = 0;
  for (int k = lo; k < hi; ++k) sum += sc[kPdCounts + k];
  part[tid] = sum;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    const int v = tid >= off ? part[tid - off] : 0;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  int run = part[tid] - sum;
  for (int k = lo; k < hi; ++k) {
    const int v = sc[kPdCounts + k];
    sc[kPdOffsets + k] = run;
    run += v;
  }
  if (tid == 255) *count = sc[kPdMiscAt] != 0 ? -1 : part[255];
}

}  // namespace b2g

// ===========================================================================
// host side
// ===========================================================================
using namespace b2g;

#define HIPCHK(expr)                                                                                  \
  do {                                                                                                \
    hipError_t e_ = (expr);                                                                           \
    if (e_ != hipSuccess) {                                                                           \
      ctx_error(ctx, std::string(#expr) + ": " + hipGetErrorString(e_));                              \
      return B2G_E_HIP;                                                                               \
    }                                                                                                 \
  } while (0)

// A run of consecutive worlds in launch order (w_order) that one launch of the island or
// pair search covers, with the dynamic LDS of its largest world.  Splitting the worlds into
// runs whose workgroups fit the same number of times into a CU's LDS keeps the many small
// worlds from being held to the footprint of the few largest; the runs of one search go to
// separate streams and run side by side.
struct WorldRun {
  int first, count;
  size_t smem;
};
constexpr int kMaxRuns = 4;
constexpr int kIslandsPerCu = 10;  // k_islands workgroups per CU allowed by its registers (128 threads)
constexpr int kIslandsWidePerCu = 5;  // ... with 256 threads
constexpr int kPairsPerCu = 7;    // k_pairs workgroups per CU allowed by its registers
constexpr int kPairsWideMaxPerCu = 2;  // runs this LDS-heavy launch 1024-thread workgroups

// Runs of `order` (at most kMaxRuns): a run ends where the next world's workgroups would
// fit more often per CU (`cap` = the kernel's limit from registers) than the run's largest;
// surplus runs are merged from the front, where the few largest worlds sit.  need[w] == 0
// marks a world the kernel skips.
static std::vector<WorldRun> plan_runs(const std::vector<int>& order, const std::vector<size_t>& need, int cap) {
  auto fits = [&](size_t s) { return s == 0 ? cap : std::min(cap, (int)((size_t)kMaxLds / s)); };
  std::vector<WorldRun> runs;
  for (int p = 0; p < (int)order.size(); ++p) {
    const size_t s = need[order[p]];
    if (runs.empty() || fits(s) > fits(runs.back().smem)) runs.push_back({p, 0, 0});
    runs.back().count += 1;
    runs.back().smem = std::max(runs.back().smem, s);
  }
  while ((int)runs.size() > kMaxRuns) {
    runs[1].first = runs[0].first;
    runs[1].count += runs[0].count;
    runs[1].smem = std::max(runs[1].smem, runs[0].smem);
    runs.erase(runs.begin());
  }
  return runs;
}

struct b2g_ctx {
  int precision = B2G_FP32;
  int device = 0;
  double contact_factor = 8.0;
  int contact_extra = 64;
  // solver launch shape (b2g_set_param)
  int large_worlds = 0;  // 1: every world takes the global-memory island and pair search
  std::vector<int> big;  // worlds on that path
  std::vector<int> toi_list;  // worlds with continuous physics (k_toi)
  int heavy_class = 5;
  int team_full_class = 9;
  int team_lds = 24576;
  int team_max = 32;
  int pairs_wide = kPairsWideMaxPerCu;   // B2G_PARAM_PAIRS_WIDE
  int islands_wide = kIslandsWidePerCu;  // B2G_PARAM_ISLANDS_WIDE
  int collide_block = kBlock;            // B2G_PARAM_COLLIDE_BLOCK
  int solve_grid = 16384;
  hipStream_t stream = nullptr;
  std::string err;
  // sizes
  int nw = 0, nb = 0, nf = 0, ncap = 0;
  int max_nb = 0, max_nf = 0, max_ccap = 0;
  size_t smem_islands = 0, smem_pairs = 0, smem_big = 0;
  std::vector<WorldRun> isl_runs, pair_runs;
  hipStream_t aux[kMaxRuns - 1] = {};  // streams of runs 1.. (run 0 stays on `stream`)
  hipEvent_t fork_ev = nullptr, join_ev[kMaxRuns - 1] = {};
  bool pending_new_fixture = false;
  size_t tree_lds = 8;      // dynamic LDS of k_tree_build
  bool tree_dirty = true;   // the fat AABBs changed since the trees were built
  int tree_pairs = 0;       // B2G_PARAM_TREE_PAIRS
  int ev_mask = 0;          // B2G_PARAM_EVENTS
  long long ev_cap = 0;     // event records the buffers hold (allocated when events are first enabled)
  // host mirrors of constant per-body / per-world data
  std::vector<int> body_off, fix_off, con_base, con_cap;
  std::vector<double> mass, inertia;
  std::vector<double> world_d;
  std::vector<int> world_i;
  std::vector<int> fix_body_global;
  // device buffers
  std::vector<void*> allocs;
  void* dev = nullptr;  // Dev<float> or Dev<double>, host copy of the pointer table
  // graph
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  double graph_dt = 0;
  int graph_vi = -1, graph_pi = -1;
  // timing
  std::vector<hipEvent_t> pev;  // b2g_profile_step
  int island_times = 0;         // B2G_PARAM_ISLAND_TIMES
  long long stat_steps = 0;     // steps launched since the work counters were reset
  long long* itime = nullptr;   // kIslandTimeFields values per island slot (nb slots)
  bool split = false;           // b2g_step_collide ran; b2g_step_solve finishes that step
  double split_dt = 0;
  int split_vi = 0, split_pi = 0;
  int* xbuf = nullptr;          // argument and result staging of the between-step calls
  size_t xcap = 0;              // ints in xbuf
  int nj = 0;                   // joints of the batch
};

static void ctx_error(b2g_ctx* ctx, const std::string& s) {
  if (ctx) ctx->err = s;
}

static int set_error(b2g_ctx* ctx, int code, const std::string& s) {
  ctx_error(ctx, s);
  return code;
}

template <typename X>
static int dalloc(b2g_ctx* ctx, X** p, size_t n) {
  *p = nullptr;
  size_t bytes = std::max<size_t>(n, 1) * sizeof(X);
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, bytes);
  if (e != hipSuccess) {
    ctx_error(ctx, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    return B2G_E_HIP;
  }
  ctx->allocs.push_back(q);
  *p = (X*)q;
  return 0;
}

static void drop_graph(b2g_ctx* ctx) {
  if (ctx->graph_exec) hipGraphExecDestroy(ctx->graph_exec);
  if (ctx->graph) hipGraphDestroy(ctx->graph);
  ctx->graph_exec = nullptr;
  ctx->graph = nullptr;
  ctx->graph_vi = -1;
}

static void free_device(b2g_ctx* ctx) {
  drop_graph(ctx);
  for (void* p : ctx->allocs) hipFree(p);
  ctx->allocs.clear();
  if (ctx->itime) hipFree(ctx->itime);
  ctx->itime = nullptr;
  if (ctx->xbuf) hipFree(ctx->xbuf);
  ctx->xbuf = nullptr;
  ctx->xcap = 0;
  if (ctx->dev) {
    if (ctx->precision == B2G_FP64) delete (Dev<double>*)ctx->dev;
    else delete (Dev<float>*)ctx->dev;
  }
  ctx->dev = nullptr;
}

template <typename X>
static int upload(b2g_ctx* ctx, X* dst, const X* src, size_t n) {
  if (n == 0) return 0;
  HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(X), hipMemcpyHostToDevice, ctx->stream));
  return 0;
}
template <typename X>
static int download(b2g_ctx* ctx, X* dst, const X* src, size_t n) {
  if (n == 0) return 0;
  HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(X), hipMemcpyDeviceToHost, ctx->stream));
  return 0;
}

#define TRY(x)           \
  do {                   \
    int rc_ = (x);       \
    if (rc_) return rc_; \
  } while (0)

// Event records for ev_cap events per step (4 per contact slot: begin/end/pre-solve in collide,
// post-solve in the solve; an end at destruction replaces the update's events).
template <typename T>
static int ensure_event_buffers(b2g_ctx* ctx, Dev<T>& d) {
  if (d.ev_i) return 0;
  const long long cap = std::min<long long>(ctx->ev_cap, 1ll << 30);
  TRY(dalloc(ctx, &d.ev_i, (size_t)cap * B2G_EV_NI));
  TRY(dalloc(ctx, &d.ev_d, (size_t)cap * B2G_EV_ND));
  d.ev_cap = (int)cap;
  return 0;
}

template <typename T>
static int upload_batch(b2g_ctx* ctx, const b2g_batch* b) {
  const int nw = b->n_worlds;
  const int nb = b->body_off[nw], nf = b->fix_off[nw], nc = b->con_off[nw];
  ctx->nw = nw;
  ctx->nb = nb;
  ctx->nf = nf;
  ctx->body_off.assign(b->body_off, b->body_off + nw + 1);
  ctx->fix_off.assign(b->fix_off, b->fix_off + nw + 1);
  ctx->con_base.resize(nw + 1);
  ctx->con_cap.resize(nw);
  ctx->max_nb = ctx->max_nf = ctx->max_ccap = 0;
  int ncap = 0;
  for (int w = 0; w < nw; ++w) {
    const int wnb = b->body_off[w + 1] - b->body_off[w];
    const int wnf = b->fix_off[w + 1] - b->fix_off[w];
    const int wnc = b->con_off[w + 1] - b->con_off[w];
    if (wnb < 0 || wnf < 0 || wnc < 0) return set_error(ctx, B2G_E_INVALID, "negative world size");
    int cap = (int)std::ceil(ctx->contact_factor * wnf) + ctx->contact_extra;
    cap = std::max(cap, 2 * wnc + 16);
    ctx->con_base[w] = ncap;
    ctx->con_cap[w] = cap;
    ncap += cap;
    ctx->max_nb = std::max(ctx->max_nb, wnb);
    ctx->max_nf = std::max(ctx->max_nf, wnf);
    ctx->max_ccap = std::max(ctx->max_ccap, cap);
    if (wnb >= 65535 || wnf >= 65535)
      return set_error(ctx, B2G_E_UNSUPPORTED, "world too large (>= 65535 bodies or fixtures)");
  }
  ctx->con_base[nw] = ncap;
  ctx->ncap = ncap;
  // worlds whose island or pair search does not fit one workgroup's LDS (or whose
  // contact slots exceed 16-bit indices) take the global-memory path
  ctx->big.clear();
  std::vector<int> w_big(nw, 0);
  ctx->smem_islands = 0;
  ctx->smem_pairs = 0;
  ctx->smem_big = 0;
  for (int w = 0; w < nw; ++w) {
    const int wnb = ctx->body_off[w + 1] - ctx->body_off[w], wnf = ctx->fix_off[w + 1] - ctx->fix_off[w];
    const size_t si = island_smem_bytes(wnb, wnf, ctx->con_cap[w]), sp = pairs_smem_bytes(wnf, ctx->con_cap[w]);
    if (ctx->large_worlds || ctx->con_cap[w] >= 65535 || si > (size_t)kMaxLds || sp > (size_t)kMaxLds) {
      w_big[w] = 1;
      ctx->big.push_back(w);
      ctx->smem_big = std::max(ctx->smem_big, (size_t)((wnb + 1) / 2) * 4);  // 16-bit union-find parents
      ctx->smem_big = std::max(ctx->smem_big, kBigSortLds);                 // level sorts, reusing them
      const size_t wl = walk_lds_bytes(wnb, ctx->con_cap[w]);              // the walks' state, if it fits
      if (wl <= kBigWalkLdsMax) ctx->smem_big = std::max(ctx->smem_big, wl);
    } else {
      ctx->smem_islands = std::max(ctx->smem_islands, si);
      ctx->smem_pairs = std::max(ctx->smem_pairs, sp);
    }
  }
  HIPCHK(hipFuncSetAttribute((const void*)k_solve<T>, hipFuncAttributeMaxDynamicSharedMemorySize, kMaxLds));
  HIPCHK(hipFuncSetAttribute((const void*)k_tree_build<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)(kTreeLdsKeys * sizeof(unsigned long long))));

  HIPCHK(hipFuncSetAttribute((const void*)k_islands<T, kBlock>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)ctx->smem_islands));
  HIPCHK(hipFuncSetAttribute((const void*)k_islands<T, kBlock / 2>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)ctx->smem_islands));
  HIPCHK(hipFuncSetAttribute((const void*)k_pairs<T, kBlock>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)ctx->smem_pairs));
  HIPCHK(hipFuncSetAttribute((const void*)k_pairs<T, kBigPairsBlock>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)ctx->smem_pairs));
  HIPCHK(hipFuncSetAttribute((const void*)k_islands_big<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)ctx->smem_big));
  Dev<T>* dp = new Dev<T>();
  ctx->dev = dp;
  Dev<T>& d = *dp;
  d.nw = nw;
  d.nb = nb;
  d.nf = nf;
  d.ncap = ncap;
  // ---- host staging ----
  std::vector<int> w_flags(nw), w_move_n(nw), w_asleep(nw, 0), w_con_used(nw), zeros_w(nw, 0);
  std::vector<double> w_grav(2 * (size_t)nw), w_inv_dt0(nw);
  ctx->world_d.assign(b->world_d, b->world_d + (size_t)nw * B2G_WORLD_ND);
  ctx->world_i.assign(b->world_i, b->world_i + (size_t)nw * B2G_WORLD_NI);
  ctx->pending_new_fixture = false;
  for (int w = 0; w < nw; ++w) {
    w_flags[w] = b->world_i[(size_t)w * B2G_WORLD_NI + B2G_WI_FLAGS];
    w_grav[2 * w] = b->world_d[(size_t)w * B2G_WORLD_ND + B2G_WD_GX];
    w_grav[2 * w + 1] = b->world_d[(size_t)w * B2G_WORLD_ND + B2G_WD_GY];
    w_inv_dt0[w] = b->world_d[(size_t)w * B2G_WORLD_ND + B2G_WD_INV_DT0];
    w_move_n[w] = b->move_off[w + 1] - b->move_off[w];
    w_con_used[w] = b->con_off[w + 1] - b->con_off[w];
    if (w_flags[w] & WF_NEW_FIXTURE) ctx->pending_new_fixture = true;
    if ((w_flags[w] & WF_CONTINUOUS) && (w_flags[w] & WF_CUSTOM_FILTER))
      return set_error(ctx, B2G_E_UNSUPPORTED,
                       "continuous physics with a custom contact filter (the filter would run inside the kernel)");
    if (w_move_n[w] > b->fix_off[w + 1] - b->fix_off[w]) return set_error(ctx, B2G_E_INVALID, "move buffer too long");
  }
  std::vector<int> b_type(nb), b_flags(nb), b_island(nb, -1);
  std::vector<double2> b_c(nb), b_c0(nb), b_lc(nb);
  std::vector<double> b_a(nb), b_a0(nb), b_sleep(nb);
  std::vector<double4> b_xf(nb);
  std::vector<V2<double>> b_v(nb);
  std::vector<V2<T>> b_f(nb);
  std::vector<double> b_w(nb);
  std::vector<T> b_im(nb), b_iI(nb), b_t(nb), b_gs(nb), b_ld(nb), b_ad(nb);
  ctx->mass.resize(nb);
  ctx->inertia.resize(nb);
  for (int w = 0; w < nw; ++w) {
    for (int i = b->body_off[w]; i < b->body_off[w + 1]; ++i) {
      const double* r = b->body_d + (size_t)i * B2G_BODY_ND;
      const int32_t* ri = b->body_i + (size_t)i * B2G_BODY_NI;
      b_type[i] = ri[B2G_BI_TYPE];
      b_flags[i] = ri[B2G_BI_FLAGS];
      if (b_type[i] < 0 || b_type[i] > 2) return set_error(ctx, B2G_E_INVALID, "bad body type");
      if (b_type[i] == BT_STATIC && (r[B2G_BD_V_X] != 0.0 || r[B2G_BD_V_Y] != 0.0 || r[B2G_BD_W] != 0.0))
        return set_error(ctx, B2G_E_UNSUPPORTED, "static body with non-zero velocity");
      if (b_type[i] != BT_STATIC && !(b_flags[i] & BF_AWAKE) && (b_flags[i] & BF_ACTIVE)) w_asleep[w]++;
      b_c[i] = make_double2(r[B2G_BD_C_X], r[B2G_BD_C_Y]);
      b_c0[i] = make_double2(r[B2G_BD_C0_X], r[B2G_BD_C0_Y]);
      b_lc[i] = make_double2(r[B2G_BD_LC_X], r[B2G_BD_LC_Y]);
      b_a[i] = r[B2G_BD_A];
      b_a0[i] = r[B2G_BD_A0];
      b_sleep[i] = r[B2G_BD_SLEEP_TIME];
      b_xf[i] = make_double4(r[B2G_BD_XF_PX], r[B2G_BD_XF_PY], r[B2G_BD_XF_C], r[B2G_BD_XF_S]);
      b_v[i] = v2<double>(r[B2G_BD_V_X], r[B2G_BD_V_Y]);
      b_w[i] = r[B2G_BD_W];
      b_f[i] = v2<T>((T)r[B2G_BD_F_X], (T)r[B2G_BD_F_Y]);
      b_t[i] = (T)r[B2G_BD_TORQUE];
      b_im[i] = (T)r[B2G_BD_INV_MASS];
      b_iI[i] = (T)r[B2G_BD_INV_I];
      b_gs[i] = (T)r[B2G_BD_GRAV_SCALE];
      b_ld[i] = (T)r[B2G_BD_LIN_DAMP];
      b_ad[i] = (T)r[B2G_BD_ANG_DAMP];
      ctx->mass[i] = r[B2G_BD_MASS];
      ctx->inertia[i] = r[B2G_BD_I];
    }
  }
  std::vector<int4> f_info(nf);
  std::vector<int2> f_wb(nf);
  std::vector<double> f_geo((size_t)nf * GEO_STRIDE, 0.0);
  std::vector<double> f_geo64((size_t)nf * 16, 0.0), f_rad64(nf);
  std::vector<V2<T>> f_mat(nf);
  std::vector<double4> f_fat(nf);
  ctx->fix_body_global.resize(nf);
  for (int w = 0; w < nw; ++w) {
    const int wnb = b->body_off[w + 1] - b->body_off[w];
    int last_body = -1;
    for (int i = b->fix_off[w]; i < b->fix_off[w + 1]; ++i) {
      const double* r = b->fix_d + (size_t)i * B2G_FIX_ND;
      const int32_t* ri = b->fix_i + (size_t)i * B2G_FIX_NI;
      const int body = ri[B2G_FI_BODY], type = ri[B2G_FI_SHAPE], nv = ri[B2G_FI_NVERTS];
      if (body < 0 || body >= wnb) return set_error(ctx, B2G_E_INVALID, "fixture body out of range");
      if (body < last_body) return set_error(ctx, B2G_E_INVALID, "fixtures must be grouped by body in body order");
      last_body = body;
      if (type == B2G_SHAPE_LOOP) return set_error(ctx, B2G_E_UNSUPPORTED, "loop shapes are not supported");
      if (type != B2G_SHAPE_CIRCLE && type != B2G_SHAPE_POLYGON && type != B2G_SHAPE_EDGE)
        return set_error(ctx, B2G_E_INVALID, "bad shape type");
      if (type == B2G_SHAPE_POLYGON && (nv < 3 || nv > B2G_MAX_POLY_VERTS))
        return set_error(ctx, B2G_E_UNSUPPORTED, "polygons need 3..8 vertices");
      if (ri[B2G_FI_SENSOR]) return set_error(ctx, B2G_E_UNSUPPORTED, "sensors are not supported");
      const int cat = ri[B2G_FI_CATEGORY], mask = ri[B2G_FI_MASK];
      if (cat < 0 || cat > 0xFFFF || mask < 0 || mask > 0xFFFF)
        return set_error(ctx, B2G_E_UNSUPPORTED, "collision bits must fit in 16 bits");
      f_info[i] = make_int4(body, type | (nv << 8) | (ri[B2G_FI_EDGE_FLAGS] << 16), cat | (mask << 16),
                            ri[B2G_FI_GROUP]);
      f_wb[i] = make_int2(w, b->body_off[w] + body);
      ctx->fix_body_global[i] = b->body_off[w] + body;
      double* g = &f_geo[(size_t)i * GEO_STRIDE];
      for (int k = 0; k < 16; ++k) g[GEO_VERTS + k] = r[B2G_FD_VERTS + k];
      for (int k = 0; k < 16; ++k) g[GEO_NORMALS + k] = r[B2G_FD_NORMALS + k];
      g[GEO_CENTROID] = r[B2G_FD_CENTROID_X];
      g[GEO_CENTROID + 1] = r[B2G_FD_CENTROID_Y];
      g[GEO_CIRCLE_P] = r[B2G_FD_CIRCLE_PX];
      g[GEO_CIRCLE_P + 1] = r[B2G_FD_CIRCLE_PY];
      for (int k = 0; k < 8; ++k) g[GEO_EDGE + k] = r[B2G_FD_EDGE_V0X + k];
      g[GEO_RADIUS] = r[B2G_FD_RADIUS];
      double* g64 = &f_geo64[(size_t)i * 16];
      if (type == B2G_SHAPE_POLYGON) {
        for (int k = 0; k < 16; ++k) g64[k] = r[B2G_FD_VERTS + k];
      } else if (type == B2G_SHAPE_CIRCLE) {
        g64[0] = r[B2G_FD_CIRCLE_PX];
        g64[1] = r[B2G_FD_CIRCLE_PY];
      } else {
        g64[0] = r[B2G_FD_EDGE_V1X];
        g64[1] = r[B2G_FD_EDGE_V1Y];
        g64[2] = r[B2G_FD_EDGE_V2X];
        g64[3] = r[B2G_FD_EDGE_V2Y];
      }
      f_rad64[i] = r[B2G_FD_RADIUS];
      f_mat[i] = v2<T>((T)r[B2G_FD_FRICTION], (T)r[B2G_FD_RESTITUTION]);
      f_fat[i] = make_double4(r[B2G_FD_FAT_LX], r[B2G_FD_FAT_LY], r[B2G_FD_FAT_UX], r[B2G_FD_FAT_UY]);
    }
  }
  std::vector<int2> c_pair(ncap, make_int2(0, 0));
  std::vector<int> c_state(ncap, 0);
  std::vector<V2<T>> c_mat(ncap);
  std::vector<V4<T>> c_lnlp(ncap), c_pts(ncap);
  std::vector<V4<double>> c_imp(ncap);
  std::vector<uint2> c_id(ncap, make_uint2(0, 0));
  for (int w = 0; w < nw; ++w) {
    const int wnf = b->fix_off[w + 1] - b->fix_off[w];
    for (int k = 0; k < b->con_off[w + 1] - b->con_off[w]; ++k) {
      const int i = b->con_off[w] + k, s = ctx->con_base[w] + k;
      const double* r = b->con_d + (size_t)i * B2G_CON_ND;
      const int32_t* ri = b->con_i + (size_t)i * B2G_CON_NI;
      if (ri[B2G_CI_FA] < 0 || ri[B2G_CI_FA] >= wnf || ri[B2G_CI_FB] < 0 || ri[B2G_CI_FB] >= wnf)
        return set_error(ctx, B2G_E_INVALID, "contact fixture out of range");
      const int fl = ri[B2G_CI_FLAGS];
      const int pc = ri[B2G_CI_PC], mt = ri[B2G_CI_MTYPE];
      if (pc < 0 || pc > 2 || mt < 0 || mt > 2) return set_error(ctx, B2G_E_INVALID, "bad manifold");
      c_pair[s] = make_int2(ri[B2G_CI_FA], ri[B2G_CI_FB]);
      c_state[s] = CS_LIVE | ((fl & B2G_CF_TOUCHING) ? CS_TOUCHING : 0) | ((fl & B2G_CF_ENABLED) ? CS_ENABLED : 0) |
                   ((fl & B2G_CF_FILTER) ? CS_FILTER : 0) | ((fl & B2G_CF_REJECT) ? CS_REJECT : 0) |
                   (mt << CS_MTYPE_SHIFT) | (pc << CS_PC_SHIFT);
      c_mat[s] = v2<T>((T)r[B2G_CD_FRICTION], (T)r[B2G_CD_RESTITUTION]);
      c_lnlp[s] = V4<T>{(T)r[B2G_CD_LN_X], (T)r[B2G_CD_LN_Y], (T)r[B2G_CD_LP_X], (T)r[B2G_CD_LP_Y]};
      c_pts[s] = V4<T>{(T)r[B2G_CD_P0_X], (T)r[B2G_CD_P0_Y], (T)r[B2G_CD_P1_X], (T)r[B2G_CD_P1_Y]};
      c_imp[s] = V4<double>{r[B2G_CD_P0_NI], r[B2G_CD_P0_TI], r[B2G_CD_P1_NI], r[B2G_CD_P1_TI]};
      c_id[s] = make_uint2((uint32_t)ri[B2G_CI_ID0], (uint32_t)ri[B2G_CI_ID1]);
    }
  }
  std::vector<int> mv(nf, 0);
  for (int w = 0; w < nw; ++w) {
    const int wnf = b->fix_off[w + 1] - b->fix_off[w];
    std::vector<char> seen(wnf, 0);
    for (int k = 0; k < b->move_off[w + 1] - b->move_off[w]; ++k) {
      const int f = b->move[b->move_off[w] + k];
      if (f < 0 || f >= wnf) return set_error(ctx, B2G_E_INVALID, "move buffer entry out of range");
      if (seen[f]) return set_error(ctx, B2G_E_INVALID, "duplicate move buffer entry");
      seen[f] = 1;
      mv[b->fix_off[w] + k] = f;
    }
  }

  // ---- device allocation ----
  int *w_body_off, *w_fix_off, *w_con_base, *w_con_cap;
  TRY(dalloc(ctx, &w_body_off, nw + 1));
  TRY(dalloc(ctx, &w_fix_off, nw + 1));
  TRY(dalloc(ctx, &w_con_base, nw + 1));
  TRY(dalloc(ctx, &w_con_cap, nw));
  d.w_body_off = w_body_off;
  d.w_fix_off = w_fix_off;
  d.w_con_base = w_con_base;
  d.w_con_cap = w_con_cap;
  TRY(dalloc(ctx, &d.w_con_used, nw));
  TRY(dalloc(ctx, &d.w_flags, nw));
  double* w_grav_d;
  TRY(dalloc(ctx, &w_grav_d, 2 * (size_t)nw));
  d.w_grav = w_grav_d;
  TRY(dalloc(ctx, &d.w_inv_dt0, nw));
  TRY(dalloc(ctx, &d.w_err, nw));
  TRY(dalloc(ctx, &d.stat, kStats));
  TRY(dalloc(ctx, &d.w_new, nw));
  TRY(dalloc(ctx, &d.ev_count, 1));
  d.ev_i = nullptr;
  d.ev_d = nullptr;
  d.ev_cap = 0;
  ctx->ev_cap = 4 * (long long)ncap + 64;
  TRY(dalloc(ctx, &d.w_move_n, nw));
  TRY(dalloc(ctx, &d.w_asleep, nw));
  TRY(dalloc(ctx, &d.w_nisl, nw));
  TRY(dalloc(ctx, &d.w_itot, nw));
  TRY(dalloc(ctx, &d.w_ipre, 2 * (nw + 1)));
  TRY(dalloc(ctx, &d.b_type, nb));
  TRY(dalloc(ctx, &d.b_flags, nb));
  TRY(dalloc(ctx, &d.b_island, nb));
  TRY(dalloc(ctx, &d.b_c, nb));
  TRY(dalloc(ctx, &d.b_a, nb));
  TRY(dalloc(ctx, &d.b_c0, nb));
  TRY(dalloc(ctx, &d.b_a0, nb));
  TRY(dalloc(ctx, &d.b_xf, nb));
  TRY(dalloc(ctx, &d.b_lc, nb));
  TRY(dalloc(ctx, &d.b_sleep, nb));
  TRY(dalloc(ctx, &d.b_v, nb));
  TRY(dalloc(ctx, &d.b_w, nb));
  TRY(dalloc(ctx, &d.b_im, nb));
  TRY(dalloc(ctx, &d.b_iI, nb));
  TRY(dalloc(ctx, &d.b_f, nb));
  TRY(dalloc(ctx, &d.b_t, nb));
  TRY(dalloc(ctx, &d.b_gs, nb));
  TRY(dalloc(ctx, &d.b_ld, nb));
  TRY(dalloc(ctx, &d.b_ad, nb));
  TRY(dalloc(ctx, &d.f_info, nf));
  TRY(dalloc(ctx, &d.f_wb, nf));
  TRY(dalloc(ctx, &d.f_geo, (size_t)nf * GEO_STRIDE));
  TRY(dalloc(ctx, &d.f_geo64, (size_t)nf * 16));
  TRY(dalloc(ctx, &d.f_rad64, nf));
  TRY(dalloc(ctx, &d.f_mat, nf));
  TRY(dalloc(ctx, &d.f_fat, nf));
  TRY(dalloc(ctx, &d.f_moved, nf));
  TRY(dalloc(ctx, &d.mv, nf));
  TRY(dalloc(ctx, &d.c_pair, ncap));
  TRY(dalloc(ctx, &d.c_state, ncap));
  TRY(dalloc(ctx, &d.c_mat, ncap));
  TRY(dalloc(ctx, &d.c_lnlp, ncap));
  TRY(dalloc(ctx, &d.c_pts, ncap));
  TRY(dalloc(ctx, &d.c_imp, ncap));
  TRY(dalloc(ctx, &d.c_id, ncap));
  TRY(dalloc(ctx, &d.i_bodies, nb));
  TRY(dalloc(ctx, &d.i_contacts, ncap));
  TRY(dalloc(ctx, &d.i_cidx, ncap));
  TRY(dalloc(ctx, &d.i_fb, ncap));
  TRY(dalloc(ctx, &d.i_isb, ncap));
  TRY(dalloc(ctx, &d.i_rank, ncap));
  TRY(dalloc(ctx, &d.i_order, ncap));
  TRY(dalloc(ctx, &d.i_info, nb));
  TRY(dalloc(ctx, &d.q_isl, nb));
  TRY(dalloc(ctx, &d.q_nc, nb));
  TRY(dalloc(ctx, &d.q_count, 1));
  TRY(dalloc(ctx, &d.qs_isl, nb));
  TRY(dalloc(ctx, &d.qs_nc, nb));
  TRY(dalloc(ctx, &d.q_cls, kQcls));
  TRY(dalloc(ctx, &d.pv, nb));
  TRY(dalloc(ctx, &d.vc, ncap));
  TRY(dalloc(ctx, &d.pc, ncap));
  std::vector<long long> hb(nw, 0), nbse(nw, 0), tsb(nw, -1);  // kept alive until the copies below complete
  std::vector<int> zero16(16 * (size_t)nw, 0);
  std::vector<int> w_order(nw);
  for (int w = 0; w < nw; ++w) w_order[w] = w;
  std::stable_sort(w_order.begin(), w_order.end(), [&](int x, int y) {
    return ctx->body_off[x + 1] - ctx->body_off[x] > ctx->body_off[y + 1] - ctx->body_off[y];
  });
  {
    int* wo = nullptr;
    TRY(dalloc(ctx, &wo, nw));
    TRY(upload(ctx, wo, w_order.data(), nw));
    d.w_order = wo;
  }
  {
    std::vector<size_t> ni(nw, 0), np(nw, 0);
    for (int w = 0; w < nw; ++w) {
      if (w_big[w]) continue;
      const int wnb = ctx->body_off[w + 1] - ctx->body_off[w], wnf = ctx->fix_off[w + 1] - ctx->fix_off[w];
      ni[w] = island_smem_bytes(wnb, wnf, ctx->con_cap[w]);
      np[w] = pairs_smem_bytes(wnf, ctx->con_cap[w]);
    }
    ctx->isl_runs = plan_runs(w_order, ni, kIslandsPerCu);
    ctx->pair_runs = plan_runs(w_order, np, kPairsPerCu);
  }
  // continuous physics: the worlds that run World.solve_toi after each step, and their state
  ctx->toi_list.clear();
  for (int w = 0; w < nw; ++w)
    if (w_flags[w] & WF_CONTINUOUS) ctx->toi_list.push_back(w);
  d.toi_worlds = nullptr;
  d.toi_scr = nullptr;
  d.toi_move = nullptr;
  d.b_alpha0 = nullptr;
  d.b_tflag = nullptr;
  d.c_toi = nullptr;
  d.c_tmeta = nullptr;
  if (!ctx->toi_list.empty()) {
    int* tl = nullptr;
    TRY(dalloc(ctx, &tl, ctx->toi_list.size()));
    TRY(upload(ctx, tl, ctx->toi_list.data(), ctx->toi_list.size()));
    d.toi_worlds = tl;
    TRY(dalloc(ctx, &d.toi_scr, ctx->toi_list.size() * sizeof(ToiScratch)));
    TRY(dalloc(ctx, &d.toi_move, nf));
    TRY(dalloc(ctx, &d.b_alpha0, nb));
    TRY(dalloc(ctx, &d.b_tflag, nb));
    TRY(dalloc(ctx, &d.c_toi, ncap));
    TRY(dalloc(ctx, &d.c_tmeta, ncap));
  }
  // joints: their records, each body's joint list (Body.joints, world.py:268-273), and the body
  // pairs a joint keeps from colliding (body.py:920-926)
  ctx->nj = 0;
  if (b->joint_off && b->joint_off[nw] > 0) {
    const int nj = b->joint_off[nw];
    if (!b->joint_d || !b->joint_i) return set_error(ctx, B2G_E_INVALID, "joint records missing");
    std::vector<double4> j_mass(4 * (size_t)nj, make_double4(0, 0, 0, 0));
    std::vector<int> bj_cnt((size_t)nb + 1, 0), jnc_off(nw + 1, 0);
    std::vector<uint32_t> jnc;
    for (int w = 0; w < nw; ++w) {
      const int j0 = b->joint_off[w], j1 = b->joint_off[w + 1], b0 = b->body_off[w];
      const int wnb = b->body_off[w + 1] - b0;
      if (j1 < j0) return set_error(ctx, B2G_E_INVALID, "negative joint count");
      if (j1 > j0 && w_big[w])
        return set_error(ctx, B2G_E_UNSUPPORTED, "joints in a world beyond the LDS island search");
      if (j1 - j0 >= 65535) return set_error(ctx, B2G_E_UNSUPPORTED, "more than 65534 joints in a world");
      std::vector<uint32_t> keys;
      for (int j = j0; j < j1; ++j) {
        const int32_t* ji = b->joint_i + (size_t)j * B2G_JOINT_NI;
        const int t = ji[B2G_JI_TYPE], ba = ji[B2G_JI_BODY_A], bb = ji[B2G_JI_BODY_B];
        if (t < B2G_JT_REVOLUTE || t > B2G_JT_ROPE) return set_error(ctx, B2G_E_INVALID, "bad joint type");
        if (bb < 0 || bb >= wnb || ba < -1 || ba >= wnb || (ba < 0 && t != B2G_JT_MOUSE))
          return set_error(ctx, B2G_E_INVALID, "joint body out of range");
        int bodies[4] = {ba, bb, -1, -1};
        if (t == B2G_JT_GEAR) {
          bodies[2] = ji[B2G_JI_BODY_C];
          bodies[3] = ji[B2G_JI_BODY_D];
          if (bodies[2] < 0 || bodies[2] >= wnb || bodies[3] < 0 || bodies[3] >= wnb)
            return set_error(ctx, B2G_E_INVALID, "gear joint body out of range");
        }
        for (int q = 0; q < 4; ++q)
          if (bodies[q] >= 0) {
            const double* r = b->body_d + (size_t)(b0 + bodies[q]) * B2G_BODY_ND;
            j_mass[4 * (size_t)j + q] = make_double4(r[B2G_BD_INV_MASS], r[B2G_BD_INV_I], r[B2G_BD_MASS], 0.0);
          }
        if (ba >= 0) bj_cnt[b0 + ba + 1]++;
        bj_cnt[b0 + bb + 1]++;
        if (!(ji[B2G_JI_FLAGS] & B2G_JF_COLLIDE_CONNECTED) && ba >= 0 && ba != bb)
          keys.push_back(ba < bb ? ((uint32_t)ba | (uint32_t)bb << 16) : ((uint32_t)bb | (uint32_t)ba << 16));
      }
      std::sort(keys.begin(), keys.end());
      keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
      jnc.insert(jnc.end(), keys.begin(), keys.end());
      jnc_off[w + 1] = (int)jnc.size();
    }
    for (int i = 0; i < nb; ++i) bj_cnt[i + 1] += bj_cnt[i];
    std::vector<uint32_t> bj(bj_cnt[nb] > 0 ? bj_cnt[nb] : 1);
    {
      std::vector<int> fill(bj_cnt.begin(), bj_cnt.end() - 1);
      for (int w = 0; w < nw; ++w) {  // world joint order gives each body's list its order
        const int j0 = b->joint_off[w], b0 = b->body_off[w];
        for (int j = j0; j < b->joint_off[w + 1]; ++j) {
          const int32_t* ji = b->joint_i + (size_t)j * B2G_JOINT_NI;
          const int ba = ji[B2G_JI_BODY_A], bb = ji[B2G_JI_BODY_B];
          const uint32_t jl = (uint32_t)(j - j0);
          if (ba >= 0) bj[fill[b0 + ba]++] = jl | (uint32_t)bb << 16;
          bj[fill[b0 + bb]++] = jl | (uint32_t)(ba >= 0 ? ba : 0xFFFF) << 16;
        }
      }
    }
    int *wjo = nullptr, *bjo = nullptr, *wjn = nullptr;
    uint32_t *bjd = nullptr, *jncd = nullptr;
    double4* jm = nullptr;
    TRY(dalloc(ctx, &wjo, nw + 1));
    TRY(upload(ctx, wjo, b->joint_off, nw + 1));
    TRY(dalloc(ctx, &d.j_d, (size_t)nj * B2G_JOINT_ND));
    TRY(upload(ctx, d.j_d, b->joint_d, (size_t)nj * B2G_JOINT_ND));
    TRY(dalloc(ctx, &d.j_i, (size_t)nj * B2G_JOINT_NI));
    TRY(upload(ctx, d.j_i, b->joint_i, (size_t)nj * B2G_JOINT_NI));
    TRY(dalloc(ctx, &jm, 4 * (size_t)nj));
    TRY(upload(ctx, jm, j_mass.data(), 4 * (size_t)nj));
    TRY(dalloc(ctx, &bjo, (size_t)nb + 1));
    TRY(upload(ctx, bjo, bj_cnt.data(), (size_t)nb + 1));
    TRY(dalloc(ctx, &bjd, bj.size()));
    TRY(upload(ctx, bjd, bj.data(), bj.size()));
    TRY(dalloc(ctx, &wjn, nw + 1));
    TRY(upload(ctx, wjn, jnc_off.data(), nw + 1));
    TRY(dalloc(ctx, &jncd, jnc.empty() ? 1 : jnc.size()));
    if (!jnc.empty()) TRY(upload(ctx, jncd, jnc.data(), jnc.size()));
    TRY(dalloc(ctx, &d.j_flag, nj));
    TRY(dalloc(ctx, &d.i_joints, nj));
    TRY(dalloc(ctx, &d.i_jidx, nj));
    TRY(dalloc(ctx, &d.i_jr, nb));
    HIPCHK(hipStreamSynchronize(ctx->stream));  // the staging vectors go out of scope
    d.w_joint_off = wjo;
    d.j_mass = jm;
    d.bj_off = bjo;
    d.bj = bjd;
    d.w_jnc_off = wjn;
    d.jnc = jncd;
    ctx->nj = nj;
  }
  if (!ctx->big.empty()) {  // scratch of the large-world path
    int *wb = nullptr, *bl = nullptr;
    TRY(dalloc(ctx, &wb, nw));
    TRY(dalloc(ctx, &bl, ctx->big.size()));
    TRY(upload(ctx, wb, w_big.data(), nw));
    TRY(upload(ctx, bl, ctx->big.data(), ctx->big.size()));
    d.w_big = wb;
    d.big_list = bl;
    TRY(dalloc(ctx, &d.g_bst, nb));
    TRY(dalloc(ctx, &d.g_wake, nb));
    TRY(dalloc(ctx, &d.g_off, (size_t)nb + nw));
    TRY(dalloc(ctx, &d.g_cur, nb));
    TRY(dalloc(ctx, &d.g_lev, nb));
    TRY(dalloc(ctx, &d.g_parent, nb));
    TRY(dalloc(ctx, &d.g_seed, nb));
    TRY(dalloc(ctx, &d.g_cntb, nb));
    TRY(dalloc(ctx, &d.g_cntc, nb));
    TRY(dalloc(ctx, &d.g_bidx, nb));
    TRY(dalloc(ctx, &d.g_adj, 2 * (size_t)ncap));
    TRY(dalloc(ctx, &d.g_cflag, ncap));
    TRY(dalloc(ctx, &d.g_stack, (size_t)nb + 2 * (size_t)ncap));
    TRY(dalloc(ctx, &d.g_isl, nb));
    TRY(dalloc(ctx, &d.g_wmisc, 16 * (size_t)nw));
    TRY(dalloc(ctx, &d.g_mpos, nf));
    TRY(dalloc(ctx, &d.g_mlist, nf));
    long long ht = 0, nt = 0;
    for (int w : ctx->big) {
      hb[w] = ht;
      nbse[w] = nt;
      ht += next_pow2(ctx->con_cap[w] + 1);
      nt += next_pow2(ctx->con_cap[w]);
    }
    long long *hbd = nullptr, *nbd = nullptr;
    TRY(dalloc(ctx, &hbd, nw));
    TRY(dalloc(ctx, &nbd, nw));
    TRY(upload(ctx, hbd, hb.data(), nw));
    TRY(upload(ctx, nbd, nbse.data(), nw));
    d.g_hbase = hbd;
    d.g_nbase = nbd;
    TRY(dalloc(ctx, &d.g_hash, (size_t)ht));
    TRY(dalloc(ctx, &d.g_nlist, (size_t)nt));
    TRY(upload(ctx, d.g_wmisc, zero16.data(), 16 * (size_t)nw));
  }
  // LBVH arrays; worlds whose keys do not fit the LDS sort get global sort scratch
  TRY(dalloc(ctx, &d.t_key, nf));
  TRY(dalloc(ctx, &d.t_child, nf));
  TRY(dalloc(ctx, &d.t_box, nf));
  TRY(dalloc(ctx, &d.t_flag, nf));
  TRY(dalloc(ctx, &d.t_root, nw));
  {
    long long st = 0;
    int lds_np = 1;
    for (int w = 0; w < nw; ++w) {
      const int np = next_pow2(std::max(1, ctx->fix_off[w + 1] - ctx->fix_off[w]));
      if (np > kTreeLdsKeys) {
        tsb[w] = st;
        st += np;
      } else {
        lds_np = std::max(lds_np, np);
      }
    }
    ctx->tree_lds = (size_t)lds_np * sizeof(unsigned long long);
    long long* tsbd = nullptr;
    TRY(dalloc(ctx, &tsbd, nw));
    TRY(upload(ctx, tsbd, tsb.data(), nw));
    d.t_sbase = tsbd;
    TRY(dalloc(ctx, &d.t_sort, (size_t)std::max<long long>(st, 1)));
    ctx->tree_dirty = true;
  }

  TRY(upload(ctx, w_body_off, b->body_off, nw + 1));
  TRY(upload(ctx, w_fix_off, b->fix_off, nw + 1));
  TRY(upload(ctx, w_con_base, ctx->con_base.data(), nw + 1));
  TRY(upload(ctx, w_con_cap, ctx->con_cap.data(), nw));
  TRY(upload(ctx, d.w_con_used, w_con_used.data(), nw));
  TRY(upload(ctx, d.w_flags, w_flags.data(), nw));
  TRY(upload(ctx, w_grav_d, w_grav.data(), 2 * (size_t)nw));
  TRY(upload(ctx, d.w_inv_dt0, w_inv_dt0.data(), nw));
  TRY(upload(ctx, d.w_err, zeros_w.data(), nw));
  TRY(upload(ctx, d.w_new, zeros_w.data(), nw));
  {
    static const int zero1 = 0;
    TRY(upload(ctx, d.ev_count, &zero1, 1));
  }
  if (ctx->ev_mask) TRY(ensure_event_buffers(ctx, d));
  {
    const std::vector<unsigned long long> z(kStats, 0ull);
    TRY(upload(ctx, d.stat, z.data(), kStats));
    ctx->stat_steps = 0;
  }
  TRY(upload(ctx, d.w_move_n, w_move_n.data(), nw));
  TRY(upload(ctx, d.w_asleep, w_asleep.data(), nw));
  TRY(upload(ctx, d.w_nisl, zeros_w.data(), nw));
  TRY(upload(ctx, d.b_type, b_type.data(), nb));
  TRY(upload(ctx, d.b_flags, b_flags.data(), nb));
  TRY(upload(ctx, d.b_island, b_island.data(), nb));
  TRY(upload(ctx, d.b_c, b_c.data(), nb));
  TRY(upload(ctx, d.b_a, b_a.data(), nb));
  TRY(upload(ctx, d.b_c0, b_c0.data(), nb));
  TRY(upload(ctx, d.b_a0, b_a0.data(), nb));
  TRY(upload(ctx, d.b_xf, b_xf.data(), nb));
  TRY(upload(ctx, d.b_lc, b_lc.data(), nb));
  TRY(upload(ctx, d.b_sleep, b_sleep.data(), nb));
  TRY(upload(ctx, d.b_v, b_v.data(), nb));
  TRY(upload(ctx, d.b_w, b_w.data(), nb));
  TRY(upload(ctx, d.b_im, b_im.data(), nb));
  TRY(upload(ctx, d.b_iI, b_iI.data(), nb));
  TRY(upload(ctx, d.b_f, b_f.data(), nb));
  TRY(upload(ctx, d.b_t, b_t.data(), nb));
  TRY(upload(ctx, d.b_gs, b_gs.data(), nb));
  TRY(upload(ctx, d.b_ld, b_ld.data(), nb));
  TRY(upload(ctx, d.b_ad, b_ad.data(), nb));
  TRY(upload(ctx, d.f_info, f_info.data(), nf));
  TRY(upload(ctx, d.f_wb, f_wb.data(), nf));
  TRY(upload(ctx, d.f_geo, f_geo.data(), (size_t)nf * GEO_STRIDE));
  TRY(upload(ctx, d.f_geo64, f_geo64.data(), (size_t)nf * 16));
  TRY(upload(ctx, d.f_rad64, f_rad64.data(), nf));
  TRY(upload(ctx, d.f_mat, f_mat.data(), nf));
  TRY(upload(ctx, d.f_fat, f_fat.data(), nf));
  TRY(upload(ctx, d.mv, mv.data(), nf));
  TRY(upload(ctx, d.c_pair, c_pair.data(), ncap));
  TRY(upload(ctx, d.c_state, c_state.data(), ncap));
  TRY(upload(ctx, d.c_mat, c_mat.data(), ncap));
  TRY(upload(ctx, d.c_lnlp, c_lnlp.data(), ncap));
  TRY(upload(ctx, d.c_pts, c_pts.data(), ncap));
  TRY(upload(ctx, d.c_imp, c_imp.data(), ncap));
  TRY(upload(ctx, d.c_id, c_id.data(), ncap));
  std::vector<int> zeros_f(nf, 0);
  TRY(upload(ctx, d.f_moved, zeros_f.data(), nf));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return 0;
}

template <typename T>
static int launch_step(b2g_ctx* ctx, double dt, int vi, int pi, bool first_pairs, hipEvent_t* ev,
                       int part = PART_ALL) {
  const bool timed = ev != nullptr;
  Dev<T> d = *(Dev<T>*)ctx->dev;
  d.dt = dt;
  d.inv_dt = dt > 0.0 ? 1.0 / dt : 0.0;
  d.vel_iters = vi;
  d.pos_iters = pi;
  d.heavy_class = ctx->heavy_class;
  d.team_full_class = ctx->team_full_class;
  d.team_lds = ctx->team_lds;
  d.team_max = ctx->team_max;
  d.ev_mask = ctx->ev_mask;
  d.solve_only = part == PART_ISLANDS;
  d.i_time = ctx->island_times ? ctx->itime : nullptr;
  const int nw = ctx->nw;
  hipStream_t st = ctx->stream;
  const int nbig = (int)ctx->big.size();
  // the runs of one search side by side: run 0 on the step's stream, the others on the
  // auxiliary streams, forked from and joined back into the step's stream
  auto fan = [&](const std::vector<WorldRun>& runs, auto&& launch) -> int {
    const int n = (int)runs.size();
    if (n > 1) HIPCHK(hipEventRecord(ctx->fork_ev, st));
    for (int r = 1; r < n; ++r) HIPCHK(hipStreamWaitEvent(ctx->aux[r - 1], ctx->fork_ev, 0));
    for (int r = 0; r < n; ++r) launch(runs[r], r == 0 ? st : ctx->aux[r - 1]);
    for (int r = 1; r < n; ++r) {
      HIPCHK(hipEventRecord(ctx->join_ev[r - 1], ctx->aux[r - 1]));
      HIPCHK(hipStreamWaitEvent(st, ctx->join_ev[r - 1], 0));
    }
    return 0;
  };
  auto pairs = [&](int from_upload, int finish, int do_pairs) -> int {
    TRY(fan(ctx->pair_runs, [&](const WorldRun& r, hipStream_t s) {
      if (r.smem > 0 && (size_t)kMaxLds / r.smem <= (size_t)ctx->pairs_wide)
        hipLaunchKernelGGL((k_pairs<T, kBigPairsBlock>), dim3(r.count), dim3(kBigPairsBlock), r.smem, s, d, r.first,
                           from_upload, finish, do_pairs);
      else
        hipLaunchKernelGGL((k_pairs<T, kBlock>), dim3(r.count), dim3(kBlock), r.smem, s, d, r.first, from_upload,
                           finish, do_pairs);
    }));
    if (nbig > 0) {
      hipLaunchKernelGGL(k_pairs_big_prep<T>, dim3(nbig), dim3(kBigBlock), 0, st, d, from_upload, do_pairs);
      if (ctx->tree_pairs && do_pairs) {  // LBVH over the large worlds' current fat AABBs
        hipLaunchKernelGGL(k_tree_build<T>, dim3(nbig), dim3(kTreeBlock), ctx->tree_lds, st, d, d.big_list);
        hipLaunchKernelGGL(k_pairs_big_find_tree<T>, dim3(64, nbig), dim3(kBlock), 0, st, d);
      } else {
        hipLaunchKernelGGL(k_pairs_big_find<T>, dim3(256, nbig), dim3(kBlock), 0, st, d);
      }
      hipLaunchKernelGGL(k_pairs_big_finish<T>, dim3(nbig), dim3(kBigBlock), 0, st, d, finish, do_pairs);
    }
    return 0;
  };
  if (part == PART_ISLANDS) {  // the island search and solve alone (b2g_solve_islands)
    hipLaunchKernelGGL(k_reset_queue<T>, dim3(1), dim3(64), 0, st, d);
  } else if (part != PART_SOLVE) {
    if (first_pairs) TRY(pairs(1, 0, 1));
    if (part == PART_PAIRS) {  // the new-fixture pair search alone (b2g_pair_search)
      HIPCHK(hipGetLastError());
      return 0;
    }
    if (timed) HIPCHK(hipEventRecord(ev[0], st));
    hipLaunchKernelGGL(k_reset_queue<T>, dim3(1), dim3(64), 0, st, d);
    {
      const int cbk = ctx->collide_block;
      const int nch = std::max((ctx->max_ccap + cbk - 1) / cbk, 1);
      const dim3 grid((unsigned)(((nw + 7) / 8) * 8 * nch));
      if (cbk == 64) hipLaunchKernelGGL((k_collide<T, 64>), grid, dim3(64), 0, st, d, nch);
      else if (cbk == 128) hipLaunchKernelGGL((k_collide<T, 128>), grid, dim3(128), 0, st, d, nch);
      else if (cbk == 512) hipLaunchKernelGGL((k_collide<T, 512>), grid, dim3(512), 0, st, d, nch);
      else hipLaunchKernelGGL((k_collide<T, 256>), grid, dim3(256), 0, st, d, nch);
    }
    if (timed) HIPCHK(hipEventRecord(ev[1], st));
    if (part == PART_COLLIDE) {  // the listener runs between collide and the solve (b2g_step_collide)
      HIPCHK(hipGetLastError());
      return 0;
    }
  }
  if (dt > 0.0) {
    TRY(fan(ctx->isl_runs, [&](const WorldRun& r, hipStream_t s) {
      if (r.smem > 0 && (size_t)kMaxLds / r.smem <= (size_t)ctx->islands_wide)
        hipLaunchKernelGGL((k_islands<T, kBlock>), dim3(r.count), dim3(kBlock), r.smem, s, d, r.first);
      else
        hipLaunchKernelGGL((k_islands<T, kBlock / 2>), dim3(r.count), dim3(kBlock / 2), r.smem, s, d, r.first);
    }));
    if (nbig > 0) {
      hipLaunchKernelGGL(k_islands_big_head<T>, dim3(nbig), dim3(kBigBlock), 0, st, d);
      hipLaunchKernelGGL(k_islands_big_ends<T>, dim3(kBigEndsGrid, nbig), dim3(kBlock), 0, st, d);
      hipLaunchKernelGGL(k_islands_big_scan<T>, dim3(nbig), dim3(kBigBlock), 0, st, d);
      hipLaunchKernelGGL(k_islands_big_fill<T>, dim3(kBigEndsGrid, nbig), dim3(kBlock), 0, st, d);
      hipLaunchKernelGGL(k_islands_big_sort<T>, dim3(kBigEndsGrid, nbig), dim3(kBlock), 0, st, d);
      hipLaunchKernelGGL(k_islands_big_seeds<T>, dim3(kBigEndsGrid, nbig), dim3(kBlock), 0, st, d);
      hipLaunchKernelGGL(k_islands_big_counts<T>, dim3(kBigEndsGrid, nbig), dim3(kBlock), 0, st, d);
      hipLaunchKernelGGL(k_islands_big<T>, dim3(nbig), dim3(kBigBlock), ctx->smem_big, st, d,
                         (int)(ctx->smem_big / (kBigBlock / 64 * 2 * sizeof(int))), (int)ctx->smem_big);
      hipLaunchKernelGGL(k_islands_big_tail<T>, dim3(kBigEndsGrid, nbig), dim3(kBlock), 0, st, d);
    }
    int nblk = std::max(1, std::min((ctx->nb + kBlock - 1) / kBlock, 1024));
    hipLaunchKernelGGL(k_order_islands<T>, dim3(nblk), dim3(kBlock), 0, st, d);
    hipLaunchKernelGGL(k_order_team_islands<T>, dim3(kSizeClasses), dim3(kPrefixBlock), 0, st, d);
    if (timed) HIPCHK(hipEventRecord(ev[2], st));
    // the island solve: per-body and per-contact setup over all islands, then the islands
    hipLaunchKernelGGL(k_island_prefix<T>, dim3(1), dim3(kPrefixBlock), 0, st, d);
    const int pgrid = std::max(1, std::min((ctx->nb + kBlock - 1) / kBlock, kPrepGrid));
    hipLaunchKernelGGL(k_prep_bodies<T>, dim3((unsigned)pgrid), dim3(kBlock), 0, st, d);
    hipLaunchKernelGGL(k_prep_contacts<T>, dim3((unsigned)kPrepGrid), dim3(kBlock), 0, st, d);
    const int swaves = std::max(1, std::min((ctx->nb + kSolveBlock - 1) / kSolveBlock, ctx->solve_grid));
    hipLaunchKernelGGL(k_solve<T>, dim3(swaves), dim3(kSolveBlock), ctx->team_lds, st, d);
    if (ctx->nj > 0)
      hipLaunchKernelGGL(k_solve_joints<T>, dim3((unsigned)std::max(1, (ctx->nb + 63) / 64)), dim3(64), 0, st, d);
    if (part == PART_ISLANDS) {
      HIPCHK(hipGetLastError());
      return 0;
    }
    if (timed) HIPCHK(hipEventRecord(ev[3], st));
    hipLaunchKernelGGL(k_sync<T>, dim3(std::max(1, (ctx->nf + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, d);
    if (timed) HIPCHK(hipEventRecord(ev[4], st));
    TRY(pairs(0, 1, 1));
    const int ntoi = (int)ctx->toi_list.size();
    if (ntoi > 0) hipLaunchKernelGGL(k_toi<T>, dim3(ntoi), dim3(64), 0, st, d, ntoi);
  } else {
    if (timed) {
      HIPCHK(hipEventRecord(ev[2], st));
      HIPCHK(hipEventRecord(ev[3], st));
      HIPCHK(hipEventRecord(ev[4], st));
    }
    TRY(pairs(0, 1, 0));
  }
  if (timed) HIPCHK(hipEventRecord(ev[5], st));
  HIPCHK(hipGetLastError());
  return 0;
}

// The island-timing buffer is allocated when timing is on and a step is about to run (the
// step entry points call this before they launch or capture anything).
static int ensure_island_times(b2g_ctx* ctx) {
  if (!ctx->island_times || ctx->itime || !ctx->dev) return 0;
  void* q = nullptr;
  const hipError_t e = hipMalloc(&q, (size_t)std::max(ctx->nb, 1) * kIslandTimeFields * sizeof(long long));
  if (e != hipSuccess) return set_error(ctx, B2G_E_HIP, std::string("hipMalloc failed: ") + hipGetErrorString(e));
  ctx->itime = (long long*)q;
  return 0;
}

static int dispatch_step(b2g_ctx* ctx, double dt, int vi, int pi, bool first_pairs, hipEvent_t* ev = nullptr,
                         int part = PART_ALL) {
  ctx->tree_dirty = true;
  if (ctx->precision == B2G_FP64) return launch_step<double>(ctx, dt, vi, pi, first_pairs, ev, part);
  return launch_step<float>(ctx, dt, vi, pi, first_pairs, ev, part);
}

template <typename T>
static int download_batch(b2g_ctx* ctx, b2g_batch_out* o) {
  Dev<T>& d = *(Dev<T>*)ctx->dev;
  const int nw = ctx->nw, nb = ctx->nb, nf = ctx->nf, ncap = ctx->ncap;
  if (o->n_worlds != nw) return set_error(ctx, B2G_E_INVALID, "n_worlds mismatch");
  std::vector<int> w_used(nw), w_flags(nw), w_move_n(nw), w_new(nw);
  std::vector<double> w_inv_dt0(nw);
  TRY(download(ctx, w_new.data(), d.w_new, nw));
  TRY(download(ctx, w_used.data(), d.w_con_used, nw));
  TRY(download(ctx, w_flags.data(), d.w_flags, nw));
  TRY(download(ctx, w_inv_dt0.data(), d.w_inv_dt0, nw));
  TRY(download(ctx, w_move_n.data(), d.w_move_n, nw));
  std::vector<int> b_type(nb), b_flags(nb), b_island(nb);
  std::vector<double2> b_c(nb), b_c0(nb), b_lc(nb);
  std::vector<double> b_a(nb), b_a0(nb), b_sleep(nb);
  std::vector<double4> b_xf(nb);
  std::vector<V2<double>> b_v(nb);
  std::vector<V2<T>> b_f(nb);
  std::vector<double> b_w(nb);
  std::vector<T> b_im(nb), b_iI(nb), b_t(nb), b_gs(nb), b_ld(nb), b_ad(nb);
  TRY(download(ctx, b_type.data(), d.b_type, nb));
  TRY(download(ctx, b_flags.data(), d.b_flags, nb));
  TRY(download(ctx, b_island.data(), d.b_island, nb));
  TRY(download(ctx, b_c.data(), d.b_c, nb));
  TRY(download(ctx, b_c0.data(), d.b_c0, nb));
  TRY(download(ctx, b_lc.data(), d.b_lc, nb));
  TRY(download(ctx, b_a.data(), d.b_a, nb));
  TRY(download(ctx, b_a0.data(), d.b_a0, nb));
  TRY(download(ctx, b_sleep.data(), d.b_sleep, nb));
  TRY(download(ctx, b_xf.data(), d.b_xf, nb));
  TRY(download(ctx, b_v.data(), d.b_v, nb));
  TRY(download(ctx, b_w.data(), d.b_w, nb));
  TRY(download(ctx, b_f.data(), d.b_f, nb));
  TRY(download(ctx, b_t.data(), d.b_t, nb));
  TRY(download(ctx, b_im.data(), d.b_im, nb));
  TRY(download(ctx, b_iI.data(), d.b_iI, nb));
  TRY(download(ctx, b_gs.data(), d.b_gs, nb));
  TRY(download(ctx, b_ld.data(), d.b_ld, nb));
  TRY(download(ctx, b_ad.data(), d.b_ad, nb));
  std::vector<double4> f_fat(nf);
  TRY(download(ctx, f_fat.data(), d.f_fat, nf));
  std::vector<int2> c_pair(ncap);
  std::vector<int> c_state(ncap);
  std::vector<V2<T>> c_mat(ncap);
  std::vector<V4<T>> c_lnlp(ncap), c_pts(ncap);
  std::vector<V4<double>> c_imp(ncap);
  std::vector<uint2> c_id(ncap);
  TRY(download(ctx, c_pair.data(), d.c_pair, ncap));
  TRY(download(ctx, c_state.data(), d.c_state, ncap));
  TRY(download(ctx, c_mat.data(), d.c_mat, ncap));
  TRY(download(ctx, c_lnlp.data(), d.c_lnlp, ncap));
  TRY(download(ctx, c_pts.data(), d.c_pts, ncap));
  TRY(download(ctx, c_imp.data(), d.c_imp, ncap));
  TRY(download(ctx, c_id.data(), d.c_id, ncap));
  std::vector<int> mv(nf);
  TRY(download(ctx, mv.data(), d.mv, nf));
  if (ctx->nj > 0 && o->joint_d && o->joint_i) {  // the joint records (joints do not change count)
    TRY(download(ctx, o->joint_d, d.j_d, (size_t)ctx->nj * B2G_JOINT_ND));
    TRY(download(ctx, o->joint_i, d.j_i, (size_t)ctx->nj * B2G_JOINT_NI));
  }
  HIPCHK(hipStreamSynchronize(ctx->stream));

  for (int w = 0; w < nw; ++w) {
    double* wd = o->world_d + (size_t)w * B2G_WORLD_ND;
    int32_t* wi = o->world_i + (size_t)w * B2G_WORLD_NI;
    for (int k = 0; k < B2G_WORLD_ND; ++k) wd[k] = ctx->world_d[(size_t)w * B2G_WORLD_ND + k];
    for (int k = 0; k < B2G_WORLD_NI; ++k) wi[k] = ctx->world_i[(size_t)w * B2G_WORLD_NI + k];
    wd[B2G_WD_INV_DT0] = w_inv_dt0[w];
    wi[B2G_WI_FLAGS] = w_flags[w];
    wi[B2G_WI_NEW_CONTACTS] = w_new[w];
  }
  for (int i = 0; i < nb; ++i) {
    double* r = o->body_d + (size_t)i * B2G_BODY_ND;
    int32_t* ri = o->body_i + (size_t)i * B2G_BODY_NI;
    std::memset(r, 0, sizeof(double) * B2G_BODY_ND);
    ri[B2G_BI_TYPE] = b_type[i];
    ri[B2G_BI_FLAGS] = b_flags[i];
    ri[B2G_BI_ISLAND] = b_island[i];
    ri[B2G_BI_PAD] = 0;
    r[B2G_BD_XF_PX] = b_xf[i].x; r[B2G_BD_XF_PY] = b_xf[i].y; r[B2G_BD_XF_C] = b_xf[i].z; r[B2G_BD_XF_S] = b_xf[i].w;
    r[B2G_BD_LC_X] = b_lc[i].x; r[B2G_BD_LC_Y] = b_lc[i].y;
    r[B2G_BD_C0_X] = b_c0[i].x; r[B2G_BD_C0_Y] = b_c0[i].y;
    r[B2G_BD_C_X] = b_c[i].x; r[B2G_BD_C_Y] = b_c[i].y;
    r[B2G_BD_A0] = b_a0[i]; r[B2G_BD_A] = b_a[i];
    r[B2G_BD_V_X] = (double)b_v[i].x; r[B2G_BD_V_Y] = (double)b_v[i].y; r[B2G_BD_W] = (double)b_w[i];
    r[B2G_BD_F_X] = (double)b_f[i].x; r[B2G_BD_F_Y] = (double)b_f[i].y; r[B2G_BD_TORQUE] = (double)b_t[i];
    r[B2G_BD_MASS] = ctx->mass[i]; r[B2G_BD_INV_MASS] = (double)b_im[i];
    r[B2G_BD_I] = ctx->inertia[i]; r[B2G_BD_INV_I] = (double)b_iI[i];
    r[B2G_BD_GRAV_SCALE] = (double)b_gs[i]; r[B2G_BD_LIN_DAMP] = (double)b_ld[i]; r[B2G_BD_ANG_DAMP] = (double)b_ad[i];
    r[B2G_BD_SLEEP_TIME] = b_sleep[i];
  }
  for (int i = 0; i < nf; ++i) {
    double* r = o->fix_d + (size_t)i * B2G_FIX_ND;
    r[B2G_FD_FAT_LX] = f_fat[i].x; r[B2G_FD_FAT_LY] = f_fat[i].y;
    r[B2G_FD_FAT_UX] = f_fat[i].z; r[B2G_FD_FAT_UY] = f_fat[i].w;
  }
  int co = 0, mo = 0;
  o->con_off[0] = 0;
  o->move_off[0] = 0;
  for (int w = 0; w < nw; ++w) {
    const int cb = ctx->con_base[w];
    for (int sl = 0; sl < w_used[w]; ++sl) {
      const int s = cb + sl;
      const int st = c_state[s];
      if (!(st & CS_LIVE)) continue;
      if (co >= o->con_cap) return set_error(ctx, B2G_E_CAPACITY, "output contact capacity too small");
      double* r = o->con_d + (size_t)co * B2G_CON_ND;
      int32_t* ri = o->con_i + (size_t)co * B2G_CON_NI;
      std::memset(r, 0, sizeof(double) * B2G_CON_ND);
      std::memset(ri, 0, sizeof(int32_t) * B2G_CON_NI);
      ri[B2G_CI_FA] = c_pair[s].x;
      ri[B2G_CI_FB] = c_pair[s].y;
      ri[B2G_CI_FLAGS] = ((st & CS_TOUCHING) ? B2G_CF_TOUCHING : 0) | ((st & CS_ENABLED) ? B2G_CF_ENABLED : 0) |
                         ((st & CS_FILTER) ? B2G_CF_FILTER : 0) | ((st & CS_REJECT) ? B2G_CF_REJECT : 0);
      ri[B2G_CI_MTYPE] = (st >> CS_MTYPE_SHIFT) & 3;
      ri[B2G_CI_PC] = (st >> CS_PC_SHIFT) & 3;
      ri[B2G_CI_ID0] = (int32_t)c_id[s].x;
      ri[B2G_CI_ID1] = (int32_t)c_id[s].y;
      r[B2G_CD_FRICTION] = (double)c_mat[s].x; r[B2G_CD_RESTITUTION] = (double)c_mat[s].y;
      r[B2G_CD_LN_X] = (double)c_lnlp[s].x; r[B2G_CD_LN_Y] = (double)c_lnlp[s].y;
      r[B2G_CD_LP_X] = (double)c_lnlp[s].z; r[B2G_CD_LP_Y] = (double)c_lnlp[s].w;
      r[B2G_CD_P0_X] = (double)c_pts[s].x; r[B2G_CD_P0_Y] = (double)c_pts[s].y;
      r[B2G_CD_P1_X] = (double)c_pts[s].z; r[B2G_CD_P1_Y] = (double)c_pts[s].w;
      r[B2G_CD_P0_NI] = (double)c_imp[s].x; r[B2G_CD_P0_TI] = (double)c_imp[s].y;
      r[B2G_CD_P1_NI] = (double)c_imp[s].z; r[B2G_CD_P1_TI] = (double)c_imp[s].w;
      ++co;
    }
    for (int k = 0; k < w_move_n[w]; ++k) {
      if (mo >= o->move_cap) return set_error(ctx, B2G_E_CAPACITY, "output move capacity too small");
      o->move[mo++] = mv[ctx->fix_off[w] + k];
    }
    o->con_off[w + 1] = co;
    o->move_off[w + 1] = mo;
  }
  return 0;
}

// ===========================================================================
// C-ABI
// ===========================================================================

// The trees of every world, rebuilt when a step or an upload changed the fat AABBs.
template <typename T>
static int ensure_trees_t(b2g_ctx* ctx) {
  if (!ctx->tree_dirty) return 0;
  const Dev<T>& d = *(Dev<T>*)ctx->dev;
  if (ctx->nw > 0)
    hipLaunchKernelGGL(k_tree_build<T>, dim3(ctx->nw), dim3(kTreeBlock), ctx->tree_lds, ctx->stream, d, d.w_order);
  HIPCHK(hipGetLastError());
  ctx->tree_dirty = false;
  return 0;
}
static int ensure_trees(b2g_ctx* ctx) {
  return ctx->precision == B2G_FP64 ? ensure_trees_t<double>(ctx) : ensure_trees_t<float>(ctx);
}

// Device staging for a query call: inputs copied in, outputs copied back, all freed on return.
struct Staging {
  std::vector<void*> p;
  ~Staging() {
    for (void* q : p) hipFree(q);
  }
  void* get(size_t bytes) {
    void* q = nullptr;
    if (hipMalloc(&q, std::max<size_t>(bytes, 16)) != hipSuccess) return nullptr;
    p.push_back(q);
    return q;
  }
};

template <typename T>
static int query_aabb_t(b2g_ctx* ctx, int n, const int32_t* world, const double* box, int32_t* out, int cap,
                        int32_t* count) {
  const Dev<T>& d = *(Dev<T>*)ctx->dev;
  Staging sg;
  int* dw = (int*)sg.get((size_t)n * 4);
  double* db = (double*)sg.get((size_t)n * 32);
  int2* dout = (int2*)sg.get((size_t)std::max(cap, 1) * 8);
  int* dcnt = (int*)sg.get(4);
  if (!dw || !db || !dout || !dcnt) return set_error(ctx, B2G_E_HIP, "hipMalloc failed (query staging)");
  TRY(upload(ctx, dw, world, (size_t)n));
  TRY(upload(ctx, db, box, 4 * (size_t)n));
  HIPCHK(hipMemsetAsync(dcnt, 0, 4, ctx->stream));
  TRY(ensure_trees(ctx));
  hipLaunchKernelGGL(k_query_aabb<T>, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, d, n, (const int*)dw,
                     (const double4*)db, dout, cap, dcnt);
  HIPCHK(hipGetLastError());
  int total = 0;
  HIPCHK(hipMemcpyAsync(&total, dcnt, 4, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  *count = total;
  const int m = std::min(total, cap);
  std::vector<int2> h((size_t)m);
  if (m > 0) HIPCHK(hipMemcpy(h.data(), dout, (size_t)m * 8, hipMemcpyDeviceToHost));
  std::sort(h.begin(), h.end(), [](int2 a, int2 b) { return a.x != b.x ? a.x < b.x : a.y < b.y; });
  for (int k = 0; k < m; ++k) {
    out[2 * k] = h[k].x;
    out[2 * k + 1] = h[k].y;
  }
  return total > cap ? set_error(ctx, B2G_E_CAPACITY, "query results exceed the output capacity") : B2G_OK;
}

template <typename T>
static int ray_cast_t(b2g_ctx* ctx, int n, const int32_t* world, const double* seg, int closest, int32_t* out_i,
                      double* out_d, int cap, int32_t* count) {
  const Dev<T>& d = *(Dev<T>*)ctx->dev;
  Staging sg;
  const int rows = closest ? n : std::max(cap, 1);
  int* dw = (int*)sg.get((size_t)n * 4);
  double* ds = (double*)sg.get((size_t)n * 32);
  int2* di = (int2*)sg.get((size_t)rows * 8);
  double* dd = (double*)sg.get((size_t)rows * 40);
  int* dcnt = (int*)sg.get(4);
  if (!dw || !ds || !di || !dd || !dcnt) return set_error(ctx, B2G_E_HIP, "hipMalloc failed (ray cast staging)");
  TRY(upload(ctx, dw, world, (size_t)n));
  TRY(upload(ctx, ds, seg, 4 * (size_t)n));
  HIPCHK(hipMemsetAsync(dcnt, 0, 4, ctx->stream));
  TRY(ensure_trees(ctx));
  hipLaunchKernelGGL(k_ray_cast<T>, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, d, n, (const int*)dw,
                     (const double4*)ds, closest, di, dd, closest ? n : cap, dcnt);
  HIPCHK(hipGetLastError());
  int total = 0;
  HIPCHK(hipMemcpyAsync(&total, dcnt, 4, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  if (closest) total = n;
  *count = total;
  const int m = std::min(total, closest ? n : cap);
  std::vector<int2> hi((size_t)m);
  std::vector<double> hd((size_t)m * 5);
  if (m > 0) {
    HIPCHK(hipMemcpy(hi.data(), di, (size_t)m * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(hd.data(), dd, (size_t)m * 40, hipMemcpyDeviceToHost));
  }
  std::vector<int> ord((size_t)m);
  for (int k = 0; k < m; ++k) ord[k] = k;
  if (!closest)  // by ray, then fraction, then fixture
    std::sort(ord.begin(), ord.end(), [&](int a, int b) {
      if (hi[a].x != hi[b].x) return hi[a].x < hi[b].x;
      if (hd[5 * a] != hd[5 * b]) return hd[5 * a] < hd[5 * b];
      return hi[a].y < hi[b].y;
    });
  for (int k = 0; k < m; ++k) {
    out_i[2 * k] = hi[ord[k]].x;
    out_i[2 * k + 1] = hi[ord[k]].y;
    for (int j = 0; j < 5; ++j) out_d[5 * k + j] = hd[5 * ord[k] + j];
  }
  return (!closest && total > cap) ? set_error(ctx, B2G_E_CAPACITY, "ray-cast hits exceed the output capacity")
                                   : B2G_OK;
}

// ---------------------------------------------------------------------------
// between-step calls of the contact listener and the custom filter
// ---------------------------------------------------------------------------
static int ensure_xbuf(b2g_ctx* ctx, size_t n) {
  if (ctx->xcap >= n) return 0;
  if (ctx->xbuf) hipFree(ctx->xbuf);
  ctx->xbuf = nullptr;
  ctx->xcap = 0;
  void* q = nullptr;
  const size_t want = std::max<size_t>(n, 1024);
  const hipError_t e = hipMalloc(&q, want * sizeof(int));
  if (e != hipSuccess) return set_error(ctx, B2G_E_HIP, std::string("hipMalloc failed: ") + hipGetErrorString(e));
  ctx->xbuf = (int*)q;
  ctx->xcap = want;
  return 0;
}

template <typename T>
static int new_contacts_t(b2g_ctx* ctx, int32_t* out, int cap, int32_t* count) {
  const Dev<T>& d = *(const Dev<T>*)ctx->dev;
  const int nw = ctx->nw;
  std::vector<int> nnew(nw), off(nw + 1, 0);
  TRY(download(ctx, nnew.data(), d.w_new, nw));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  for (int w = 0; w < nw; ++w) off[w + 1] = off[w] + nnew[w];
  const int total = off[nw];
  *count = total;
  if (total > cap) return set_error(ctx, B2G_E_CAPACITY, "b2g_new_contacts: output capacity too small");
  if (total == 0) return B2G_OK;
  TRY(ensure_xbuf(ctx, (size_t)(nw + 1) + 4 * (size_t)total));
  int* d_off = ctx->xbuf;
  int* d_out = ctx->xbuf + (nw + 1);
  TRY(upload(ctx, d_off, off.data(), nw + 1));
  hipLaunchKernelGGL(k_new_contacts<T>, dim3((total + kBlock - 1) / kBlock), dim3(kBlock), 0, ctx->stream, d, d_off,
                     total, d_out);
  HIPCHK(hipGetLastError());
  TRY(download(ctx, out, d_out, 4 * (size_t)total));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return B2G_OK;
}

template <typename T>
static int drop_contacts_t(b2g_ctx* ctx, int n, const int32_t* world, const int32_t* index) {
  const Dev<T>& d = *(const Dev<T>*)ctx->dev;
  const int nw = ctx->nw;
  std::vector<int> nnew(nw);
  TRY(download(ctx, nnew.data(), d.w_new, nw));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  std::vector<int> wj(2 * (size_t)n);
  for (int i = 0; i < n; ++i) {
    if (world[i] < 0 || world[i] >= nw) return set_error(ctx, B2G_E_INVALID, "b2g_drop_contacts: bad world index");
    if (index[i] < 0 || index[i] >= nnew[world[i]])
      return set_error(ctx, B2G_E_INVALID, "b2g_drop_contacts: not a contact of the world's last pair search");
    wj[2 * i] = world[i];
    wj[2 * i + 1] = index[i];
  }
  TRY(ensure_xbuf(ctx, 2 * (size_t)n));
  TRY(upload(ctx, ctx->xbuf, wj.data(), wj.size()));
  hipLaunchKernelGGL(k_drop_new<T>, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, ctx->stream, d, ctx->xbuf, n);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(ctx->stream));  // the staging buffer is reused by the next call
  ctx->tree_dirty = true;
  return B2G_OK;
}

template <typename T>
static int disable_contacts_t(b2g_ctx* ctx, int n, const int32_t* world, const int32_t* slot) {
  const Dev<T>& d = *(const Dev<T>*)ctx->dev;
  const int nw = ctx->nw;
  std::vector<int> used(nw);
  TRY(download(ctx, used.data(), d.w_con_used, nw));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  std::vector<int> ws(2 * (size_t)n);
  for (int i = 0; i < n; ++i) {
    if (world[i] < 0 || world[i] >= nw) return set_error(ctx, B2G_E_INVALID, "b2g_disable_contacts: bad world index");
    if (slot[i] < 0 || slot[i] >= used[world[i]])
      return set_error(ctx, B2G_E_INVALID, "b2g_disable_contacts: no such contact slot");
    ws[2 * i] = world[i];
    ws[2 * i + 1] = slot[i];
  }
  TRY(ensure_xbuf(ctx, 2 * (size_t)n));
  TRY(upload(ctx, ctx->xbuf, ws.data(), ws.size()));
  hipLaunchKernelGGL(k_disable<T>, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, ctx->stream, d, ctx->xbuf, n);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return B2G_OK;
}

template <typename T>
static int events_t(b2g_ctx* ctx, int32_t* out_i, double* out_d, int cap, int32_t* count) {
  const Dev<T>& d = *(const Dev<T>*)ctx->dev;
  int n = 0;
  if (d.ev_i) {
    TRY(download(ctx, &n, d.ev_count, 1));
    HIPCHK(hipStreamSynchronize(ctx->stream));
  }
  if (n > d.ev_cap) return set_error(ctx, B2G_E_CAPACITY, "b2g_events: event buffer overflow");
  *count = n;
  if (n > cap) return set_error(ctx, B2G_E_CAPACITY, "b2g_events: output capacity too small");
  if (n == 0) return B2G_OK;
  std::vector<int32_t> ei((size_t)n * B2G_EV_NI);
  std::vector<double> ed((size_t)n * B2G_EV_ND);
  TRY(download(ctx, ei.data(), d.ev_i, ei.size()));
  TRY(download(ctx, ed.data(), d.ev_d, ed.size()));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  // records arrive in completion order; the listener's call order is (world, phase, key)
  std::vector<int> idx(n);
  for (int k = 0; k < n; ++k) idx[k] = k;
  auto key = [&](int k) {
    const int32_t* r = &ei[(size_t)k * B2G_EV_NI];
    return std::make_tuple(r[B2G_EI_WORLD], r[B2G_EI_PHASE], r[B2G_EI_KEY]);
  };
  std::sort(idx.begin(), idx.end(), [&](int a, int b) { return key(a) < key(b); });
  for (int k = 0; k < n; ++k) {
    std::memcpy(out_i + (size_t)k * B2G_EV_NI, &ei[(size_t)idx[k] * B2G_EV_NI], sizeof(int32_t) * B2G_EV_NI);
    std::memcpy(out_d + (size_t)k * B2G_EV_ND, &ed[(size_t)idx[k] * B2G_EV_ND], sizeof(double) * B2G_EV_ND);
  }
  return B2G_OK;
}

extern "C" {

B2G_API int b2g_version(void) { return B2G_VERSION; }

static void destroy_streams(b2g_ctx* ctx) {
  for (int r = 0; r < kMaxRuns - 1; ++r) {
    if (ctx->join_ev[r]) hipEventDestroy(ctx->join_ev[r]);
    if (ctx->aux[r]) hipStreamDestroy(ctx->aux[r]);
  }
  if (ctx->fork_ev) hipEventDestroy(ctx->fork_ev);
  if (ctx->stream) hipStreamDestroy(ctx->stream);
}

B2G_API int b2g_create(const b2g_desc* desc, b2g_ctx** out) {
  if (!out) return B2G_E_INVALID;
  *out = nullptr;
  b2g_ctx* ctx = new b2g_ctx();
  if (desc) {
    if (desc->precision != B2G_FP32 && desc->precision != B2G_FP64) {
      delete ctx;
      return B2G_E_INVALID;
    }
    ctx->precision = desc->precision;
    ctx->device = desc->device;
    if (desc->contact_factor > 0) ctx->contact_factor = desc->contact_factor;
    if (desc->contact_extra > 0) ctx->contact_extra = desc->contact_extra;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    delete ctx;
    return B2G_E_NODEVICE;
  }
  if (ctx->device < 0 || ctx->device >= ndev || hipSetDevice(ctx->device) != hipSuccess) {
    delete ctx;
    return B2G_E_NODEVICE;
  }
  bool ok = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) == hipSuccess &&
            hipEventCreateWithFlags(&ctx->fork_ev, hipEventDisableTiming) == hipSuccess;
  for (int r = 0; ok && r < kMaxRuns - 1; ++r)
    ok = hipStreamCreateWithFlags(&ctx->aux[r], hipStreamNonBlocking) == hipSuccess &&
         hipEventCreateWithFlags(&ctx->join_ev[r], hipEventDisableTiming) == hipSuccess;
  if (!ok) {
    destroy_streams(ctx);
    delete ctx;
    return B2G_E_HIP;
  }

  *out = ctx;
  return B2G_OK;
}

B2G_API int b2g_destroy(b2g_ctx* ctx) {
  if (!ctx) return B2G_OK;
  hipSetDevice(ctx->device);
  if (ctx->stream) hipStreamSynchronize(ctx->stream);
  free_device(ctx);
  for (auto& e : ctx->pev) hipEventDestroy(e);
  ctx->pev.clear();
  destroy_streams(ctx);
  delete ctx;
  return B2G_OK;
}

B2G_API int b2g_upload(b2g_ctx* ctx, const b2g_batch* batch) {
  if (!ctx || !batch || batch->n_worlds <= 0) return set_error(ctx, B2G_E_INVALID, "bad arguments");
  hipSetDevice(ctx->device);
  hipStreamSynchronize(ctx->stream);
  free_device(ctx);
  ctx->split = false;
  int rc = ctx->precision == B2G_FP64 ? upload_batch<double>(ctx, batch) : upload_batch<float>(ctx, batch);
  if (rc) free_device(ctx);
  return rc;
}

B2G_API int b2g_step(b2g_ctx* ctx, double dt, int vel_iters, int pos_iters) {
  if (!ctx || !ctx->dev) return set_error(ctx, B2G_E_INVALID, "no batch uploaded");
  if (ctx->split) return set_error(ctx, B2G_E_INVALID, "a step begun by b2g_step_collide is not finished");
  if (vel_iters < 0 || pos_iters < 0) return set_error(ctx, B2G_E_INVALID, "negative iteration count");
  hipSetDevice(ctx->device);
  TRY(ensure_island_times(ctx));
  bool first = ctx->pending_new_fixture;
  ctx->pending_new_fixture = false;
  TRY(dispatch_step(ctx, dt, vel_iters, pos_iters, first));
  ctx->stat_steps += 1;
  return B2G_OK;
}

B2G_API int b2g_step_n(b2g_ctx* ctx, int n, double dt, int vel_iters, int pos_iters, int use_graph) {
  if (!ctx || !ctx->dev) return set_error(ctx, B2G_E_INVALID, "no batch uploaded");
  if (ctx->split) return set_error(ctx, B2G_E_INVALID, "a step begun by b2g_step_collide is not finished");
  if (n < 0 || vel_iters < 0 || pos_iters < 0) return set_error(ctx, B2G_E_INVALID, "negative count");
  hipSetDevice(ctx->device);
  TRY(ensure_island_times(ctx));  // before any capture
  int i = 0;
  if (ctx->pending_new_fixture && n > 0) {
    ctx->pending_new_fixture = false;
    TRY(dispatch_step(ctx, dt, vel_iters, pos_iters, true));
    ++i;
  }
  ctx->stat_steps += n;
  if (!use_graph) {
    for (; i < n; ++i) TRY(dispatch_step(ctx, dt, vel_iters, pos_iters, false));
    return B2G_OK;
  }
  if (i < n) {
    if (!ctx->graph_exec || ctx->graph_dt != dt || ctx->graph_vi != vel_iters || ctx->graph_pi != pos_iters) {
      if (ctx->graph_exec) hipGraphExecDestroy(ctx->graph_exec);
      if (ctx->graph) hipGraphDestroy(ctx->graph);
      ctx->graph_exec = nullptr;
      ctx->graph = nullptr;
      HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
      int rc = dispatch_step(ctx, dt, vel_iters, pos_iters, false);
      hipGraph_t g = nullptr;
      hipError_t e = hipStreamEndCapture(ctx->stream, &g);
      if (rc) return rc;
      if (e != hipSuccess) return set_error(ctx, B2G_E_HIP, std::string("capture: ") + hipGetErrorString(e));
      ctx->graph = g;
      HIPCHK(hipGraphInstantiate(&ctx->graph_exec, g, nullptr, nullptr, 0));
      ctx->graph_dt = dt;
      ctx->graph_vi = vel_iters;
      ctx->graph_pi = pos_iters;
    }
    for (; i < n; ++i) HIPCHK(hipGraphLaunch(ctx->graph_exec, ctx->stream));
  }
  return B2G_OK;
}

B2G_API int b2g_sync(b2g_ctx* ctx) {
  if (!ctx) return B2G_E_INVALID;
  hipSetDevice(ctx->device);
  HIPCHK(hipStreamSynchronize(ctx->stream));
  if (!ctx->dev) return B2G_OK;
  std::vector<int> err(ctx->nw);
  int* w_err = ctx->precision == B2G_FP64 ? ((Dev<double>*)ctx->dev)->w_err : ((Dev<float>*)ctx->dev)->w_err;
  HIPCHK(hipMemcpy(err.data(), w_err, sizeof(int) * ctx->nw, hipMemcpyDeviceToHost));
  int e = 0;
  for (int v : err) e |= v;
  if (e & WERR_CAPACITY) return set_error(ctx, B2G_E_CAPACITY, "contact segment capacity exceeded");
  if (e & WERR_NEWPAIRS) return set_error(ctx, B2G_E_CAPACITY, "too many new pairs in one step");
  if (e & WERR_UNSUPPORTED) return set_error(ctx, B2G_E_UNSUPPORTED, "unsupported shape pair");
  if (e & WERR_NAN) return set_error(ctx, B2G_E_NAN, "non-finite body position, angle or velocity");
  if (e & WERR_ISLAND_SIZE)
    return set_error(ctx, B2G_E_UNSUPPORTED, "an island of more than 65535 contacts (not solved)");
  if (e & WERR_SCHEDULE) return set_error(ctx, B2G_E_HIP, "island solver pass limit reached (internal error)");
  if (e & WERR_TOI)
    return set_error(ctx, B2G_E_UNSUPPORTED,
                     "a time-of-impact search did not end (the reference's push-back loop has no bound)");
  if (e & WERR_JOINT)
    return set_error(ctx, B2G_E_INVALID,
                     "a mouse joint's spring is too weak for the time step (the reference asserts, mouse.py:173)");
  return B2G_OK;
}

B2G_API int b2g_counts(b2g_ctx* ctx, int32_t* counts) {
  if (!ctx || !ctx->dev || !counts) return B2G_E_INVALID;
  hipSetDevice(ctx->device);
  TRY(b2g_sync(ctx));
  const int nw = ctx->nw;
  std::vector<int> used(nw), isl(nw);
  int* d_used = ctx->precision == B2G_FP64 ? ((Dev<double>*)ctx->dev)->w_con_used : ((Dev<float>*)ctx->dev)->w_con_used;
  int* d_isl = ctx->precision == B2G_FP64 ? ((Dev<double>*)ctx->dev)->w_nisl : ((Dev<float>*)ctx->dev)->w_nisl;
  HIPCHK(hipMemcpy(used.data(), d_used, sizeof(int) * nw, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(isl.data(), d_isl, sizeof(int) * nw, hipMemcpyDeviceToHost));
  long long su = 0, si = 0;
  for (int w = 0; w < nw; ++w) {
    su += used[w];
    si += isl[w];
  }
  counts[0] = ctx->nb;
  counts[1] = ctx->nf;
  counts[2] = (int32_t)std::min<long long>(su, 0x7FFFFFFF);  // slots in use (live contacts are <= this)
  counts[3] = (int32_t)std::min<long long>(si, 0x7FFFFFFF);  // islands of the last step
  return B2G_OK;
}

B2G_API int b2g_download(b2g_ctx* ctx, b2g_batch_out* out) {
  if (!ctx || !ctx->dev || !out) return set_error(ctx, B2G_E_INVALID, "bad arguments");
  hipSetDevice(ctx->device);
  TRY(b2g_sync(ctx));
  return ctx->precision == B2G_FP64 ? download_batch<double>(ctx, out) : download_batch<float>(ctx, out);
}

B2G_API int b2g_last_error(b2g_ctx* ctx, char* buf, size_t n) {
  if (!ctx || !buf || n == 0) return B2G_E_INVALID;
  std::snprintf(buf, n, "%s", ctx->err.c_str());
  return B2G_OK;
}

B2G_API int b2g_set_param(b2g_ctx* ctx, int32_t key, int32_t value) {
  if (!ctx) return B2G_E_INVALID;
  int* slot = nullptr;
  int lo = 0, hi = 0;
  switch (key) {
    case B2G_PARAM_HEAVY_CLASS: slot = &ctx->heavy_class; lo = 0; hi = kSizeClasses; break;
    case B2G_PARAM_TEAM_FULL_CLASS: slot = &ctx->team_full_class; lo = 0; hi = kSizeClasses; break;
    case B2G_PARAM_LARGE_WORLDS: slot = &ctx->large_worlds; lo = 0; hi = 1; break;
    case B2G_PARAM_TEAM_LDS: slot = &ctx->team_lds; lo = 0; hi = kMaxLds; break;
    case B2G_PARAM_SOLVE_GRID: slot = &ctx->solve_grid; lo = 1; hi = 1 << 30; break;
    case B2G_PARAM_ISLAND_TIMES: slot = &ctx->island_times; lo = 0; hi = 1; break;
    case B2G_PARAM_TEAM_MAX: slot = &ctx->team_max; lo = 8; hi = kSolveBlock; break;
    case B2G_PARAM_PAIRS_WIDE: slot = &ctx->pairs_wide; lo = 0; hi = 32; break;
    case B2G_PARAM_ISLANDS_WIDE: slot = &ctx->islands_wide; lo = 0; hi = 32; break;
    case B2G_PARAM_COLLIDE_BLOCK: slot = &ctx->collide_block; lo = 64; hi = 512; break;
    case B2G_PARAM_TREE_PAIRS: slot = &ctx->tree_pairs; lo = 0; hi = 1; break;
    case B2G_PARAM_EVENTS: slot = &ctx->ev_mask; lo = 0; hi = 15; break;
    case B2G_PARAM_SOLVE_WAVES_PER_CU:
      return set_error(ctx, B2G_E_INVALID, "b2g_set_param: B2G_PARAM_SOLVE_WAVES_PER_CU is read-only");
    case B2G_PARAM_SEARCH_RUNS:
      return set_error(ctx, B2G_E_INVALID, "b2g_set_param: B2G_PARAM_SEARCH_RUNS is read-only");
    default: return set_error(ctx, B2G_E_INVALID, "b2g_set_param: unknown key " + std::to_string(key));
  }
  if (key == B2G_PARAM_COLLIDE_BLOCK && (value & (value - 1)) != 0)
    return set_error(ctx, B2G_E_INVALID, "b2g_set_param: the collide workgroup must be 64, 128, 256 or 512 slots");
  if (key == B2G_PARAM_TEAM_MAX && (value & (value - 1)) != 0)
    return set_error(ctx, B2G_E_INVALID, "b2g_set_param: the largest team must be 8, 16, 32 or 64 lanes");
  if (key == B2G_PARAM_TEAM_LDS && value % 256 != 0)
    return set_error(ctx, B2G_E_INVALID, "b2g_set_param: the team LDS budget must be a multiple of 256");
  if (value < lo || value > hi)
    return set_error(ctx, B2G_E_INVALID, "b2g_set_param: value " + std::to_string(value) + " outside [" +
                                             std::to_string(lo) + ", " + std::to_string(hi) + "]");
  if (*slot != value) {
    if (ctx->stream) {
      hipSetDevice(ctx->device);
      HIPCHK(hipStreamSynchronize(ctx->stream));
    }
    *slot = value;
    drop_graph(ctx);  // kernel arguments are baked into a captured step
  }
  if (key == B2G_PARAM_EVENTS && value && ctx->dev) {
    hipSetDevice(ctx->device);
    TRY(ctx->precision == B2G_FP64 ? ensure_event_buffers(ctx, *(Dev<double>*)ctx->dev)
                                   : ensure_event_buffers(ctx, *(Dev<float>*)ctx->dev));
  }
  return B2G_OK;
}

B2G_API int b2g_get_param(b2g_ctx* ctx, int32_t key, int32_t* value) {
  if (!ctx || !value) return B2G_E_INVALID;
  switch (key) {
    case B2G_PARAM_HEAVY_CLASS: *value = ctx->heavy_class; return B2G_OK;
    case B2G_PARAM_TEAM_FULL_CLASS: *value = ctx->team_full_class; return B2G_OK;
    case B2G_PARAM_LARGE_WORLDS: *value = ctx->large_worlds; return B2G_OK;
    case B2G_PARAM_TEAM_LDS: *value = ctx->team_lds; return B2G_OK;
    case B2G_PARAM_SOLVE_GRID: *value = ctx->solve_grid; return B2G_OK;
    case B2G_PARAM_ISLAND_TIMES: *value = ctx->island_times; return B2G_OK;
    case B2G_PARAM_TEAM_MAX: *value = ctx->team_max; return B2G_OK;
    case B2G_PARAM_PAIRS_WIDE: *value = ctx->pairs_wide; return B2G_OK;
    case B2G_PARAM_ISLANDS_WIDE: *value = ctx->islands_wide; return B2G_OK;
    case B2G_PARAM_COLLIDE_BLOCK: *value = ctx->collide_block; return B2G_OK;
    case B2G_PARAM_TREE_PAIRS: *value = ctx->tree_pairs; return B2G_OK;
    case B2G_PARAM_EVENTS: *value = ctx->ev_mask; return B2G_OK;
    case B2G_PARAM_SEARCH_RUNS: *value = (int)ctx->isl_runs.size() | (int)ctx->pair_runs.size() << 8; return B2G_OK;
    case B2G_PARAM_SOLVE_WAVES_PER_CU: {
      // resident solver wavefronts per compute unit: 4 SIMDs x (512 VGPRs / VGPRs per lane), and
      // the LDS the launch reserves per one-wavefront workgroup
      hipSetDevice(ctx->device);
      hipFuncAttributes fa;
      const void* fn = ctx->precision == B2G_FP64 ? (const void*)k_solve<double> : (const void*)k_solve<float>;
      HIPCHK(hipFuncGetAttributes(&fa, fn));
      const int vgpr = fa.numRegs > 0 ? ((fa.numRegs + 7) / 8) * 8 : 512;
      const int by_vgpr = 4 * std::min(8, 512 / vgpr);
      const int by_lds = ctx->team_lds > 0 ? kMaxLds / ctx->team_lds : 32;
      *value = std::min(by_vgpr, by_lds);
      return B2G_OK;
    }
    default: return set_error(ctx, B2G_E_INVALID, "b2g_get_param: unknown key " + std::to_string(key));
  }
}

B2G_API int b2g_profile_step(b2g_ctx* ctx, int n, double dt, int vel_iters, int pos_iters, float* ms_out) {
  if (!ctx || !ctx->dev || !ms_out || n <= 0) return set_error(ctx, B2G_E_INVALID, "bad arguments");
  if (ctx->split) return set_error(ctx, B2G_E_INVALID, "a step begun by b2g_step_collide is not finished");
  hipSetDevice(ctx->device);
  TRY(ensure_island_times(ctx));
  // a pending new-fixture pair search runs at the start of the first step, before its first
  // event (as b2g_step does), so exactly n steps run
  bool first = ctx->pending_new_fixture;
  ctx->pending_new_fixture = false;
  if (ctx->pev.empty()) {  // a ring of 6 events per step, read once per batch of steps
    ctx->pev.resize(6 * kProfileBatch);
    for (auto& e : ctx->pev) HIPCHK(hipEventCreate(&e));
  }
  double acc[5] = {0, 0, 0, 0, 0};
  for (int done = 0; done < n;) {
    const int m = std::min(kProfileBatch, n - done);
    for (int i = 0; i < m; ++i) {
      TRY(dispatch_step(ctx, dt, vel_iters, pos_iters, first, &ctx->pev[6 * i]));
      first = false;
    }
    HIPCHK(hipEventSynchronize(ctx->pev[6 * (m - 1) + 5]));
    for (int i = 0; i < m; ++i)
      for (int k = 0; k < 5; ++k) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, ctx->pev[6 * i + k], ctx->pev[6 * i + k + 1]));
        acc[k] += ms;
      }
    done += m;
  }
  for (int k = 0; k < 5; ++k) ms_out[k] = (float)(acc[k] / n);
  ctx->stat_steps += n;
  return B2G_OK;
}

B2G_API int b2g_query_aabb(b2g_ctx* ctx, int n, const int32_t* world, const double* box, int32_t* out, int cap,
                           int32_t* count) {
  if (!ctx || !ctx->dev) return set_error(ctx, B2G_E_INVALID, "no batch uploaded");
  if (n < 0 || cap < 0 || !count || (n > 0 && (!world || !box)) || (cap > 0 && !out))
    return set_error(ctx, B2G_E_INVALID, "b2g_query_aabb: bad arguments");
  *count = 0;
  for (int k = 0; k < n; ++k)
    if (world[k] < 0 || world[k] >= ctx->nw) return set_error(ctx, B2G_E_INVALID, "b2g_query_aabb: bad world index");
  if (n == 0) return B2G_OK;
  hipSetDevice(ctx->device);
  return ctx->precision == B2G_FP64 ? query_aabb_t<double>(ctx, n, world, box, out, cap, count)
                                    : query_aabb_t<float>(ctx, n, world, box, out, cap, count);
}

B2G_API int b2g_ray_cast(b2g_ctx* ctx, int n, const int32_t* world, const double* seg, int32_t closest, int32_t* out_i,
                         double* out_d, int cap, int32_t* count) {
  if (!ctx || !ctx->dev) return set_error(ctx, B2G_E_INVALID, "no batch uploaded");
  if (n < 0 || cap < 0 || !count || (n > 0 && (!world || !seg)) || ((closest ? n : cap) > 0 && (!out_i || !out_d)))
    return set_error(ctx, B2G_E_INVALID, "b2g_ray_cast: bad arguments");
  if (closest && cap < n) return set_error(ctx, B2G_E_INVALID, "b2g_ray_cast: closest hits need cap >= n");
  *count = 0;
  for (int k = 0; k < n; ++k) {
    if (world[k] < 0 || world[k] >= ctx->nw) return set_error(ctx, B2G_E_INVALID, "b2g_ray_cast: bad world index");
    const double rx = seg[4 * k + 2] - seg[4 * k], ry = seg[4 * k + 3] - seg[4 * k + 1];
    if (!(rx * rx + ry * ry > 0.0))  // DynamicTree.ray_cast raises ValueError (dynamictree.py:225-227)
      return set_error(ctx, B2G_E_INVALID, "b2g_ray_cast: ||p2 - p1||^2 must be > 0");
  }
  if (n == 0) return B2G_OK;
  hipSetDevice(ctx->device);
  return ctx->precision == B2G_FP64 ? ray_cast_t<double>(ctx, n, world, seg, closest, out_i, out_d, cap, count)
                                    : ray_cast_t<float>(ctx, n, world, seg, closest, out_i, out_d, cap, count);
}

B2G_API int b2g_pair_search(b2g_ctx* ctx) {
  if (!ctx || !ctx->dev) return set_error(ctx, B2G_E_INVALID, "no batch uploaded");
  if (ctx->split) return set_error(ctx, B2G_E_INVALID, "a step begun by b2g_step_collide is not finished");
  hipSetDevice(ctx->device);
  if (!ctx->pending_new_fixture) {  // no search pending: nothing is new
    const int* w_new = ctx->precision == B2G_FP64 ? ((Dev<double>*)ctx->dev)->w_new : ((Dev<float>*)ctx->dev)->w_new;
    HIPCHK(hipMemsetAsync((void*)w_new, 0, sizeof(int) * ctx->nw, ctx->stream));
    return B2G_OK;
  }
  ctx->pending_new_fixture = false;
  TRY(dispatch_step(ctx, 0.0, 0, 0, true, nullptr, PART_PAIRS));
  const dim3 grid((unsigned)((ctx->nw + kBlock - 1) / kBlock));
  if (ctx->precision == B2G_FP64)
    hipLaunchKernelGGL(k_clear_new_fixture<double>, grid, dim3(kBlock), 0, ctx->stream, *(Dev<double>*)ctx->dev);
  else
    hipLaunchKernelGGL(k_clear_new_fixture<float>, grid, dim3(kBlock), 0, ctx->stream, *(Dev<float>*)ctx->dev);
  HIPCHK(hipGetLastError());
  return B2G_OK;
}

B2G_API int b2g_step_collide(b2g_ctx* ctx, double dt, int vel_iters, int pos_iters) {
  if (!ctx || !ctx->dev) return set_error(ctx, B2G_E_INVALID, "no batch uploaded");
  if (vel_iters < 0 || pos_iters < 0) return set_error(ctx, B2G_E_INVALID, "negative iteration count");
  if (ctx->split) return set_error(ctx, B2G_E_INVALID, "b2g_step_collide: the last split step was not finished");
  hipSetDevice(ctx->device);
  TRY(ensure_island_times(ctx));
  const bool first = ctx->pending_new_fixture;
  ctx->pending_new_fixture = false;
  TRY(dispatch_step(ctx, dt, vel_iters, pos_iters, first, nullptr, PART_COLLIDE));
  ctx->split = true;
  ctx->split_dt = dt;
  ctx->split_vi = vel_iters;
  ctx->split_pi = pos_iters;
  return B2G_OK;
}

B2G_API int b2g_step_solve(b2g_ctx* ctx) {
  if (!ctx || !ctx->dev) return set_error(ctx, B2G_E_INVALID, "no batch uploaded");
  if (!ctx->split) return set_error(ctx, B2G_E_INVALID, "b2g_step_solve: no step begun with b2g_step_collide");
  hipSetDevice(ctx->device);
  ctx->split = false;
  TRY(dispatch_step(ctx, ctx->split_dt, ctx->split_vi, ctx->split_pi, false, nullptr, PART_SOLVE));
  ctx->stat_steps += 1;
  return B2G_OK;
}

B2G_API int b2g_solve_islands(b2g_ctx* ctx, double dt, int vel_iters, int pos_iters) {
  if (!ctx || !ctx->dev) return set_error(ctx, B2G_E_INVALID, "no batch uploaded");
  if (ctx->split) return set_error(ctx, B2G_E_INVALID, "a step begun by b2g_step_collide is not finished");
  if (vel_iters < 0 || pos_iters < 0) return set_error(ctx, B2G_E_INVALID, "negative iteration count");
  if (!(dt > 0.0)) return set_error(ctx, B2G_E_INVALID, "the island solve needs dt > 0");
  hipSetDevice(ctx->device);
  TRY(ensure_island_times(ctx));
  TRY(dispatch_step(ctx, dt, vel_iters, pos_iters, false, nullptr, PART_ISLANDS));
  return B2G_OK;
}

B2G_API int b2g_disable_contacts(b2g_ctx* ctx, int32_t n, const int32_t* world, const int32_t* slot) {
  if (!ctx || !ctx->dev) return set_error(ctx, B2G_E_INVALID, "no batch uploaded");
  if (n < 0 || (n > 0 && (!world || !slot))) return set_error(ctx, B2G_E_INVALID, "b2g_disable_contacts: bad arguments");
  if (!ctx->split) return set_error(ctx, B2G_E_INVALID, "b2g_disable_contacts: only between collide and solve");
  if (n == 0) return B2G_OK;
  hipSetDevice(ctx->device);
  TRY(b2g_sync(ctx));
  return ctx->precision == B2G_FP64 ? disable_contacts_t<double>(ctx, n, world, slot)
                                    : disable_contacts_t<float>(ctx, n, world, slot);
}

B2G_API int b2g_new_contacts(b2g_ctx* ctx, int32_t* out, int32_t cap, int32_t* count) {
  if (!ctx || !ctx->dev) return set_error(ctx, B2G_E_INVALID, "no batch uploaded");
  if (!count || cap < 0 || (cap > 0 && !out)) return set_error(ctx, B2G_E_INVALID, "b2g_new_contacts: bad arguments");
  *count = 0;
  hipSetDevice(ctx->device);
  TRY(b2g_sync(ctx));  // a step that ran out of contact slots listed nothing usable
  return ctx->precision == B2G_FP64 ? new_contacts_t<double>(ctx, out, cap, count)
                                    : new_contacts_t<float>(ctx, out, cap, count);
}

B2G_API int b2g_drop_contacts(b2g_ctx* ctx, int32_t n, const int32_t* world, const int32_t* index) {
  if (!ctx || !ctx->dev) return set_error(ctx, B2G_E_INVALID, "no batch uploaded");
  if (n < 0 || (n > 0 && (!world || !index))) return set_error(ctx, B2G_E_INVALID, "b2g_drop_contacts: bad arguments");
  if (n == 0) return B2G_OK;
  hipSetDevice(ctx->device);
  TRY(b2g_sync(ctx));
  return ctx->precision == B2G_FP64 ? drop_contacts_t<double>(ctx, n, world, index)
                                    : drop_contacts_t<float>(ctx, n, world, index);
}

B2G_API int b2g_events(b2g_ctx* ctx, int32_t* out_i, double* out_d, int32_t cap, int32_t* count) {
  if (!ctx || !ctx->dev) return set_error(ctx, B2G_E_INVALID, "no batch uploaded");
  if (!count || cap < 0 || (cap > 0 && (!out_i || !out_d)))
    return set_error(ctx, B2G_E_INVALID, "b2g_events: bad arguments");
  *count = 0;
  hipSetDevice(ctx->device);
  TRY(b2g_sync(ctx));
  return ctx->precision == B2G_FP64 ? events_t<double>(ctx, out_i, out_d, cap, count)
                                    : events_t<float>(ctx, out_i, out_d, cap, count);
}

B2G_API int b2g_stats(b2g_ctx* ctx, int64_t* out, int32_t n, int32_t reset) {
  if (!ctx || !ctx->dev || !out || n < 1) return set_error(ctx, B2G_E_INVALID, "bad arguments");
  hipSetDevice(ctx->device);
  HIPCHK(hipStreamSynchronize(ctx->stream));
  unsigned long long* st = ctx->precision == B2G_FP64 ? ((Dev<double>*)ctx->dev)->stat : ((Dev<float>*)ctx->dev)->stat;
  unsigned long long v[kStats];
  HIPCHK(hipMemcpy(v, st, sizeof(v), hipMemcpyDeviceToHost));
  out[0] = ctx->stat_steps;
  for (int k = 0; k < 8 && k + 1 < n; ++k) out[k + 1] = (int64_t)v[k];
  if (reset) {
    HIPCHK(hipMemset(st, 0, sizeof(v)));
    ctx->stat_steps = 0;
  }
  return B2G_OK;
}

B2G_API int b2g_island_times(b2g_ctx* ctx, int64_t* out, int32_t cap, int32_t* count) {
  if (!ctx || !ctx->dev || !count || cap < 0 || (cap > 0 && !out)) return set_error(ctx, B2G_E_INVALID, "bad arguments");
  if (!ctx->island_times || !ctx->itime)
    return set_error(ctx, B2G_E_INVALID, "island timing is off (set B2G_PARAM_ISLAND_TIMES, then step)");
  hipSetDevice(ctx->device);
  HIPCHK(hipStreamSynchronize(ctx->stream));
  const int* qc = ctx->precision == B2G_FP64 ? ((Dev<double>*)ctx->dev)->q_count : ((Dev<float>*)ctx->dev)->q_count;
  int n = 0;
  HIPCHK(hipMemcpy(&n, qc, sizeof(int), hipMemcpyDeviceToHost));
  *count = n;
  const int m = std::min(n, (int)cap);
  if (m <= 0) return B2G_OK;
  int rate_khz = 0;
  HIPCHK(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, ctx->device));
  if (rate_khz <= 0) return set_error(ctx, B2G_E_HIP, "no wall clock rate");
  constexpr int F = kIslandTimeFields;
  std::vector<long long> raw((size_t)F * m);
  HIPCHK(hipMemcpy(raw.data(), ctx->itime, raw.size() * sizeof(long long), hipMemcpyDeviceToHost));
  for (int i = 0; i < m; ++i) {
    const long long* r = &raw[(size_t)F * i];
    int64_t* o = out + (size_t)F * i;
    o[0] = (int64_t)((double)r[0] * 1.0e6 / rate_khz);
    o[1] = (int64_t)((double)r[1] * 1.0e6 / rate_khz);
    o[2] = r[2];
    o[3] = r[3];
    o[4] = (int64_t)((double)r[4] * 1.0e6 / rate_khz);
    o[5] = r[5];
  }
  return B2G_OK;
}

B2G_API int b2g_narrowphase(int precision, int n, const double* fa_d, const int32_t* fa_i, const double* fb_d,
                            const int32_t* fb_i, const double* xfa, const double* xfb, double* con_d,
                            int32_t* con_i) {
  if (n < 0 || (precision != B2G_FP32 && precision != B2G_FP64)) return B2G_E_INVALID;
  if (n == 0) return B2G_OK;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return B2G_E_NODEVICE;
  const bool f64 = precision == B2G_FP64;
  const size_t ts = f64 ? sizeof(double) : sizeof(float);
  std::vector<double> ga((size_t)n * GEO_STRIDE), gb((size_t)n * GEO_STRIDE);
  std::vector<int> ia(n), ib(n);
  auto fill = [&](const double* fd, const int32_t* fi, std::vector<double>& g, std::vector<int>& info) {
    for (int k = 0; k < n; ++k) {
      const double* r = fd + (size_t)k * B2G_FIX_ND;
      const int32_t* ri = fi + (size_t)k * B2G_FIX_NI;
      double tmp[GEO_STRIDE] = {0};
      for (int j = 0; j < 16; ++j) tmp[GEO_VERTS + j] = r[B2G_FD_VERTS + j];
      for (int j = 0; j < 16; ++j) tmp[GEO_NORMALS + j] = r[B2G_FD_NORMALS + j];
      tmp[GEO_CENTROID] = r[B2G_FD_CENTROID_X];
      tmp[GEO_CENTROID + 1] = r[B2G_FD_CENTROID_Y];
      tmp[GEO_CIRCLE_P] = r[B2G_FD_CIRCLE_PX];
      tmp[GEO_CIRCLE_P + 1] = r[B2G_FD_CIRCLE_PY];
      for (int j = 0; j < 8; ++j) tmp[GEO_EDGE + j] = r[B2G_FD_EDGE_V0X + j];
      tmp[GEO_RADIUS] = r[B2G_FD_RADIUS];
      for (int j = 0; j < GEO_STRIDE; ++j) g[(size_t)k * GEO_STRIDE + j] = tmp[j];
      info[k] = ri[B2G_FI_SHAPE] | (ri[B2G_FI_NVERTS] << 8) | (ri[B2G_FI_EDGE_FLAGS] << 16);
    }
  };
  fill(fa_d, fa_i, ga, ia);
  fill(fb_d, fb_i, gb, ib);
  void *dga = nullptr, *dgb = nullptr, *dia = nullptr, *dib = nullptr, *dxa = nullptr, *dxb = nullptr,
       *doi = nullptr, *dot_ = nullptr;
  int rc = B2G_OK;
  hipStream_t st = nullptr;
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return B2G_E_HIP;
  if (hipMalloc(&dga, ga.size() * 8) != hipSuccess || hipMalloc(&dgb, gb.size() * 8) != hipSuccess ||
      hipMalloc(&dia, n * 4) != hipSuccess || hipMalloc(&dib, n * 4) != hipSuccess ||
      hipMalloc(&dxa, (size_t)n * 32) != hipSuccess || hipMalloc(&dxb, (size_t)n * 32) != hipSuccess ||
      hipMalloc(&doi, (size_t)n * 16) != hipSuccess || hipMalloc(&dot_, (size_t)n * 8 * ts) != hipSuccess) {
    rc = B2G_E_HIP;
  }
  std::vector<int> oi((size_t)n * 4);
  std::vector<unsigned char> ot((size_t)n * 8 * ts);
  if (!rc) {
    hipMemcpyAsync(dga, ga.data(), ga.size() * 8, hipMemcpyHostToDevice, st);
    hipMemcpyAsync(dgb, gb.data(), gb.size() * 8, hipMemcpyHostToDevice, st);
    hipMemcpyAsync(dia, ia.data(), n * 4, hipMemcpyHostToDevice, st);
    hipMemcpyAsync(dib, ib.data(), n * 4, hipMemcpyHostToDevice, st);
    hipMemcpyAsync(dxa, xfa, (size_t)n * 32, hipMemcpyHostToDevice, st);
    hipMemcpyAsync(dxb, xfb, (size_t)n * 32, hipMemcpyHostToDevice, st);
    const int blocks = (n + 127) / 128;
    if (f64)
      hipLaunchKernelGGL(k_narrow_seam<double>, dim3(blocks), dim3(128), 0, st, n, (const double*)dga,
                         (const int*)dia, (const double*)dgb, (const int*)dib, (const double*)dxa,
                         (const double*)dxb, (int*)doi, (double*)dot_);
    else
      hipLaunchKernelGGL(k_narrow_seam<float>, dim3(blocks), dim3(128), 0, st, n, (const double*)dga,
                         (const int*)dia, (const double*)dgb, (const int*)dib, (const double*)dxa,
                         (const double*)dxb, (int*)doi, (float*)dot_);
    hipMemcpyAsync(oi.data(), doi, (size_t)n * 16, hipMemcpyDeviceToHost, st);
    hipMemcpyAsync(ot.data(), dot_, (size_t)n * 8 * ts, hipMemcpyDeviceToHost, st);
    if (hipStreamSynchronize(st) != hipSuccess) rc = B2G_E_HIP;
  }
  hipFree(dga); hipFree(dgb); hipFree(dia); hipFree(dib); hipFree(dxa); hipFree(dxb); hipFree(doi); hipFree(dot_);
  hipStreamDestroy(st);
  if (rc) return rc;
  for (int k = 0; k < n; ++k) {
    double* r = con_d + (size_t)k * B2G_CON_ND;
    int32_t* ri = con_i + (size_t)k * B2G_CON_NI;
    std::memset(r, 0, sizeof(double) * B2G_CON_ND);
    std::memset(ri, 0, sizeof(int32_t) * B2G_CON_NI);
    if (oi[4 * k] < 0) {
      rc = B2G_E_UNSUPPORTED;
      continue;
    }
    ri[B2G_CI_MTYPE] = oi[4 * k];
    ri[B2G_CI_PC] = oi[4 * k + 1];
    ri[B2G_CI_ID0] = oi[4 * k + 2];
    ri[B2G_CI_ID1] = oi[4 * k + 3];
    double v[8];
    for (int j = 0; j < 8; ++j)
      v[j] = f64 ? ((double*)ot.data())[(size_t)k * 8 + j] : (double)((float*)ot.data())[(size_t)k * 8 + j];
    r[B2G_CD_LN_X] = v[0]; r[B2G_CD_LN_Y] = v[1]; r[B2G_CD_LP_X] = v[2]; r[B2G_CD_LP_Y] = v[3];
    r[B2G_CD_P0_X] = v[4]; r[B2G_CD_P0_Y] = v[5]; r[B2G_CD_P1_X] = v[6]; r[B2G_CD_P1_Y] = v[7];
  }
  return rc;
}

B2G_API int b2g_pairs(int n, const double* fat, int nmove, const int32_t* move, int32_t* out, int cap,
                      int32_t* count) {
  if (n < 0 || nmove < 0 || !count || (n > 0 && !fat) || (nmove > 0 && !move)) return B2G_E_INVALID;
  *count = 0;
  if (n == 0 || nmove == 0) return B2G_OK;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return B2G_E_NODEVICE;
  std::vector<int> mpos(n, -1);
  for (int k = 0; k < nmove; ++k) {
    if (move[k] < 0 || move[k] >= n) return B2G_E_INVALID;
    if (mpos[move[k]] >= 0) return B2G_E_INVALID;  // duplicate entries are not part of this seam
    mpos[move[k]] = k;
  }
  const long long work = (long long)nmove * n;
  const int dcap = (int)std::min<long long>(work, 1 << 26);
  void *dfat = nullptr, *dmove = nullptr, *dmpos = nullptr, *dout = nullptr, *dcnt = nullptr;
  int rc = B2G_OK;
  hipStream_t st = nullptr;
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return B2G_E_HIP;
  if (hipMalloc(&dfat, (size_t)n * 32) != hipSuccess || hipMalloc(&dmove, (size_t)nmove * 4) != hipSuccess ||
      hipMalloc(&dmpos, (size_t)n * 4) != hipSuccess || hipMalloc(&dout, (size_t)dcap * 8) != hipSuccess ||
      hipMalloc(&dcnt, 4) != hipSuccess)
    rc = B2G_E_HIP;
  int hcount = 0;
  std::vector<uint32_t> hout;
  if (!rc) {
    hipMemcpyAsync(dfat, fat, (size_t)n * 32, hipMemcpyHostToDevice, st);
    hipMemcpyAsync(dmove, move, (size_t)nmove * 4, hipMemcpyHostToDevice, st);
    hipMemcpyAsync(dmpos, mpos.data(), (size_t)n * 4, hipMemcpyHostToDevice, st);
    int zero = 0;
    hipMemcpyAsync(dcnt, &zero, 4, hipMemcpyHostToDevice, st);
    const long long blocks = (work + 255) / 256;
    hipLaunchKernelGGL(k_pairs_seam, dim3((unsigned)blocks), dim3(256), 0, st, n, (const double4*)dfat, nmove,
                       (const int*)dmove, (const int*)dmpos, (uint32_t*)dout, dcap, (int*)dcnt);
    hipMemcpyAsync(&hcount, dcnt, 4, hipMemcpyDeviceToHost, st);
    if (hipStreamSynchronize(st) != hipSuccess) rc = B2G_E_HIP;
    if (!rc) {
      if (hcount > dcap) rc = B2G_E_CAPACITY;
      else {
        hout.resize((size_t)hcount * 2);
        hipMemcpy(hout.data(), dout, (size_t)hcount * 8, hipMemcpyDeviceToHost);
      }
    }
  }
  hipFree(dfat); hipFree(dmove); hipFree(dmpos); hipFree(dout); hipFree(dcnt);
  hipStreamDestroy(st);
  if (rc) return rc;
  // canonical order: move-buffer position, then partner index
  std::vector<uint64_t> keys(hcount);
  for (int k = 0; k < hcount; ++k) keys[k] = ((uint64_t)hout[2 * k] << 32) | hout[2 * k + 1];
  std::sort(keys.begin(), keys.end());
  if (hcount > cap) {
    *count = hcount;
    return B2G_E_CAPACITY;
  }
  for (int k = 0; k < hcount; ++k) {
    out[2 * k] = move[(int)(keys[k] >> 32)];
    out[2 * k + 1] = (int32_t)(keys[k] & 0xFFFFFFFFu);
  }
  *count = hcount;
  return B2G_OK;
}

// Per device: the address of the pair seam's scratch array (hipGetSymbolAddress, taken once
// per device with that device current, at the entry point's first call there).
static std::mutex g_pd_mu;
static int* g_pd_scratch[64] = {};

static int pd_scratch(int dev, int** out) {
  if (dev < 0 || dev >= 64) return B2G_E_NODEVICE;
  std::lock_guard<std::mutex> lk(g_pd_mu);
  if (!g_pd_scratch[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return B2G_E_HIP;
    g_pd_scratch[dev] = (int*)p;
  }
  *out = g_pd_scratch[dev];
  return B2G_OK;
}

B2G_API int b2g_pairs_noalloc(int n, const double* fat, int nmove, const int32_t* move, int32_t* out, int cap,
                              int32_t* count, void* stream) {
  if (n < 0 || nmove < 0 || cap < 0 || !count || (n > 0 && !fat) || (nmove > 0 && !move) || (cap > 0 && !out))
    return B2G_E_INVALID;
  if (n > kPdMaxProxies || nmove > kPdMaxMoved) return B2G_E_INVALID;  // beyond the scratch array
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess) return B2G_E_NODEVICE;
  int* sc = nullptr;
  const int rc = pd_scratch(dev, &sc);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  const int fill = std::max(1, std::min((n + 255) / 256, 4096));
  hipLaunchKernelGGL(k_pd_init, dim3(fill), dim3(256), 0, st, n, sc);
  if (nmove > 0)
    hipLaunchKernelGGL(k_pd_mark, dim3(std::min((nmove + 255) / 256, 4096)), dim3(256), 0, st, n, nmove,
                       (const int*)move, sc);
  if (nmove > 0 && n > 0)
    hipLaunchKernelGGL((k_pd_pairs<false>), dim3(nmove), dim3(256), 0, st, n, (const double4*)fat, nmove,
                       (const int*)move, (const int*)sc, sc, (int*)out, cap);
  hipLaunchKernelGGL(k_pd_scan, dim3(1), dim3(256), 0, st, nmove, sc, (int*)count);
  if (nmove > 0 && n > 0)
    hipLaunchKernelGGL((k_pd_pairs<true>), dim3(nmove), dim3(256), 0, st, n, (const double4*)fat, nmove,
                       (const int*)move, (const int*)sc, sc, (int*)out, cap);
  return hipGetLastError() == hipSuccess ? B2G_OK : B2G_E_HIP;
}

}  // extern "C"
