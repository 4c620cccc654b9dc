/*
 * b2g.h -- C-ABI of the MI355X contact-pipeline library (libb2g.so).
 *
 * The reference, pypybox2d, has no native boundary on this path: the per-step
 * contact pipeline sits behind Python method seams.  Each entry point below
 * names the reference interface it replaces.  A Python caller binds them with
 * ctypes (pypybox2d_amd/_native.py); the data records are described in
 * b2g_layout.h.
 *
 * All functions return B2G_OK (0) or a negative status.  A context owns one
 * HIP stream on one GPU and is used from one host thread at a time.  No
 * callback crosses the boundary; the caller owns every host array it passes
 * and the library copies what it keeps.
 */
#ifndef B2G_H
#define B2G_H

#include <stddef.h>
#include <stdint.h>

#include "b2g_layout.h"

#ifdef __cplusplus
extern "C" {
#endif

#define B2G_API __attribute__((visibility("default")))
#define B2G_VERSION 1

/* status codes */
#define B2G_OK 0
#define B2G_E_INVALID (-1)     /* bad argument or record (reference: ValueError) */
#define B2G_E_CAPACITY (-2)    /* a device buffer is too small; re-create with larger capacity */
#define B2G_E_HIP (-3)         /* a HIP runtime call failed */
#define B2G_E_NAN (-4)         /* non-finite state detected */
#define B2G_E_UNSUPPORTED (-5) /* feature outside this path (sensor, loop shape, ...) */
#define B2G_E_NODEVICE (-6)    /* no usable GPU */

/* bulk precision of a context */
#define B2G_FP32 0 /* FP32 manifolds/solver; FP64 positions, angles, fat AABBs, sleep timers */
#define B2G_FP64 1 /* everything FP64 */

typedef struct b2g_desc {
  int32_t precision;     /* B2G_FP32 or B2G_FP64 */
  int32_t device;        /* HIP device ordinal */
  double contact_factor; /* contact slots per fixture in each world (default 8) */
  int32_t contact_extra; /* extra contact slots per world (default 64) */
  int32_t reserved;
} b2g_desc;

typedef struct b2g_ctx b2g_ctx;

B2G_API int b2g_version(void);

/* Create / destroy a context (one GPU; one stream for the step and three
   helper streams on which the island and pair searches run side by side,
   forked from and joined back into the step's stream).  Replaces the construction
   of World (world.py:64-91) and its ContactManager/BroadPhase
   (contact_manager.py:34-44, broadphase.py:36-41) as far as the device state
   goes. */
B2G_API int b2g_create(const b2g_desc* desc, b2g_ctx** out);
B2G_API int b2g_destroy(b2g_ctx* ctx);

/* Copy a batch of worlds to the device (replaces the world-building calls
   World.create_body / Body.attach_fixture / Fixture._create_proxies,
   world.py:183-233, body.py:185-214, fixture.py:250-261, and World.add_joint,
   world.py:252-283, whose results the caller encodes in the records).
   Replaces any previous batch. */
B2G_API int b2g_upload(b2g_ctx* ctx, const b2g_batch* batch);

/* Advance every world by one World.step(dt, vel_iters, pos_iters)
   (world.py:307-354): ContactManager.collide (contact_manager.py:174-223),
   World._solve with Island.solve and its joints (world.py:475-592,
   island.py:65-179, joints/), find_new_contacts (contact_manager.py:102-158)
   and, for worlds with B2G_WF_CONTINUOUS, World.solve_toi (world.py:594-821).
   Asynchronous; see b2g_sync. */
B2G_API int b2g_step(b2g_ctx* ctx, double dt, int vel_iters, int pos_iters);

/* n steps; with use_graph != 0 the step is captured once in a HIP graph and
   replayed (same semantics as n calls of b2g_step). */
B2G_API int b2g_step_n(b2g_ctx* ctx, int n, double dt, int vel_iters, int pos_iters, int use_graph);

/* Wait for queued work and report device-side errors (capacity, unsupported). */
B2G_API int b2g_sync(b2g_ctx* ctx);

/* counts[0..3] = bodies, fixtures, contact slots in use, islands of the last step. */
B2G_API int b2g_counts(b2g_ctx* ctx, int32_t* counts);

/* Copy the full state back (bodies, fat AABBs, ordered contacts with their
   manifolds and impulses, move buffer, joint records) into caller arrays.
   Replaces reading Body/Contact/Manifold/Joint attributes after World.step. */
B2G_API int b2g_download(b2g_ctx* ctx, b2g_batch_out* out);

B2G_API int b2g_last_error(b2g_ctx* ctx, char* buf, size_t n);

/* Launch shape of the island solver.  The reference solves islands one after
   another on one thread (world.py:475-580, island.py:65-179), so these have no
   counterpart there; they change only how islands are dealt to wavefronts,
   never the order of work inside an island, and so never the results.
     B2G_PARAM_HEAVY_CLASS      an island is heavy when its size class
                                floor(log2(contacts + 1)) is at least this,
                                0..16 (16: none); a heavy island is solved by
                                a team of lanes, the others by one lane each
     B2G_PARAM_TEAM_FULL_CLASS  heavy islands of this class and above get a
                                whole wavefront (64 lanes), one class below 32
                                lanes, then 16, and 8 for the rest; 0..16
     B2G_PARAM_TEAM_MAX         the largest team, 8, 16, 32 or 64 lanes: a
                                wavefront carries 64 / team islands (default 32)
     B2G_PARAM_TEAM_LDS         LDS bytes per wavefront, shared by its teams,
                                a multiple of 256 up to 163840; an island whose
                                state needs more than its team's share (24 B
                                per body and 4 per body rounded to even
                                bodies) gets a whole wavefront and its budget,
                                and if it needs more than that, one lane
     B2G_PARAM_SOLVE_GRID       most wavefronts launched for the solver, >= 1
     B2G_PARAM_LARGE_WORLDS     1: every world takes the global-memory island
                                and pair search (otherwise only worlds whose
                                search does not fit 160 KB of LDS, or with
                                >= 65535 contact slots); read at b2g_upload
     B2G_PARAM_ISLAND_TIMES     1: the solver records each island's start and
                                end time (b2g_island_times); 0 (default): off
     B2G_PARAM_SOLVE_WAVES_PER_CU  read only: solver wavefronts resident per
                                compute unit under the current launch shape
                                (the lesser of the VGPR and LDS limits)
     B2G_PARAM_SEARCH_RUNS      read only: launches per step of the island
                                search (low byte) and of the pair search (next
                                byte); each covers a run of worlds of similar
                                LDS footprint, set at b2g_upload, up to 4 */
#define B2G_PARAM_TEAM_LDS 1
#define B2G_PARAM_HEAVY_CLASS 2
#define B2G_PARAM_SOLVE_GRID 3
#define B2G_PARAM_TEAM_FULL_CLASS 4
#define B2G_PARAM_LARGE_WORLDS 5
#define B2G_PARAM_ISLAND_TIMES 6
#define B2G_PARAM_SOLVE_WAVES_PER_CU 7
#define B2G_PARAM_TEAM_MAX 8
#define B2G_PARAM_SEARCH_RUNS 9
/* Launch shapes of the searches and of the narrow phase (results never change):
     B2G_PARAM_PAIRS_WIDE       pair-search runs whose LDS footprint lets at most
                                this many workgroups share a CU launch
                                1024-thread workgroups, the others 256 (0..32,
                                default 2)
     B2G_PARAM_ISLANDS_WIDE     island-search runs whose LDS footprint lets at
                                most this many workgroups share a CU launch
                                256-thread workgroups, the others 128 (0..32,
                                default 5)
     B2G_PARAM_COLLIDE_BLOCK    contact slots per narrow-phase workgroup: 64,
                                128, 256 (default) or 512 */
#define B2G_PARAM_PAIRS_WIDE 10
#define B2G_PARAM_ISLANDS_WIDE 11
#define B2G_PARAM_COLLIDE_BLOCK 12
/* B2G_PARAM_TREE_PAIRS   1: the large worlds' pair search (worlds on the
                          global-memory path) traverses an LBVH built over the
                          world's fat AABBs after each synchronisation instead of
                          testing every moved proxy against every proxy; 0
                          (default): the all-pairs test.  Same pairs, same order. */
#define B2G_PARAM_TREE_PAIRS 13
/* B2G_PARAM_EVENTS       the contact listener's callbacks to record, a mask of
                          1 << B2G_EV_BEGIN / END / PRE_SOLVE / POST_SOLVE
                          (0, default: none).  Records are appended by the step's
                          kernels and read with b2g_events; the buffers (4 records
                          of 224 bytes per contact slot) are allocated when a mask
                          is first set. */
#define B2G_PARAM_EVENTS 14
B2G_API int b2g_set_param(b2g_ctx* ctx, int32_t key, int32_t value);
B2G_API int b2g_get_param(b2g_ctx* ctx, int32_t key, int32_t* value);

/* Run n eager steps with HIP events between the step's kernels and write the
   average milliseconds per stage: ms_out[0..4] = collide (with the queue
   reset), islands (island search and the size-ordered queue), solve (the
   per-body and per-contact setup kernels and the island solve), sync, pairs.
   The events are read back once per 128 steps. */
B2G_API int b2g_profile_step(b2g_ctx* ctx, int n, double dt, int vel_iters, int pos_iters, float* ms_out);

/* With B2G_PARAM_ISLAND_TIMES on: the islands the last step solved, in the
   solver's queue order, as 6 int64 each: start and end (ns, device wall
   clock), contacts | bodies << 32, lanes | period << 32 (the lanes that
   solved it: a team's size, 1 for one lane per island, 0 for a team island
   whose state did not fit the team's LDS share or that has joints; and its
   sweep period, the lag in dependency levels between consecutive sweeps,
   which with the iteration counts predicts the island's passes), the end of
   its velocity iterations (ns) and the number of position iterations it ran
   before the early exit.  *count receives the number of islands; at most
   cap rows are written.  No counterpart in the reference (per-island latency
   of the parallel solver). */
B2G_API int b2g_island_times(b2g_ctx* ctx, int64_t* out, int32_t cap, int32_t* count);

/* World.query_aabb (world.py:371-384) for n boxes at once: for query k, box
   box[4k .. 4k+3] = (lower x, lower y, upper x, upper y) in world world[k],
   every fixture whose fat AABB overlaps it (DynamicTree.query,
   dynamictree.py:190-207, with PyAABB.test_overlap), as out[2j] = k,
   out[2j + 1] = world-local fixture index, sorted by query then fixture.  The
   tree is an LBVH built on the device over the state of the last step.
   *count receives the number of results; with more than cap, the first cap are
   written and B2G_E_CAPACITY is returned. */
B2G_API int b2g_query_aabb(b2g_ctx* ctx, int n, const int32_t* world, const double* box, int32_t* out, int cap,
                           int32_t* count);

/* World.ray_cast (world.py:386-440) for n segments at once: seg[4k .. 4k+3] =
   (p1 x, p1 y, p2 x, p2 y) in world world[k].  A fixture is hit when its proxy
   passes DynamicTree.ray_cast's leaf test (dynamictree.py:209-273) and its shape's
   ray cast (shapes.py:147-174, 265-307, 655-702) hits with max_fraction 1.
   Rows: out_i[2j] = ray, out_i[2j + 1] = world-local fixture (-1: no hit);
   out_d[5j ..] = fraction, point x, y ((1 - f) p1 + f p2), normal x, y.
   closest != 0: row k is ray k's hit of least fraction (ties: lower fixture
   index), cap >= n.  closest == 0: every hit, sorted by ray, fraction, fixture;
   more than cap hits: the first cap are written, B2G_E_CAPACITY.  A segment
   with p1 == p2 is rejected (B2G_E_INVALID; the reference raises ValueError). */
B2G_API int b2g_ray_cast(b2g_ctx* ctx, int n, const int32_t* world, const double* seg, int32_t closest, int32_t* out_i,
                         double* out_d, int cap, int32_t* count);

/* Work counters of the steps since the last reset (summed over steps):
   out[0] steps, out[1] live contacts at the narrow phase (fat-AABB pairs present
   in the step: ContactManager.collide's loop, contact_manager.py:174-223),
   out[2] of those evaluated (an awake non-static body), out[3] contacts created
   (find_new_contacts, contact_manager.py:102-158), out[4] proxies re-fattened
   and moved (DynamicTree.move_proxy returning true, dynamictree.py:152-188),
   out[5] island bodies and out[6] island contacts solved (the lengths of
   Island.bodies and Island.contacts, world.py:494-580, summed over islands),
   with continuous physics out[7] time_of_impact calls (toi.py:215) and out[8]
   TOI events (sub-step island solves, world.py:801).
   At most n values are written; reset != 0 zeroes the counters.  No
   counterpart in the reference (the per-step counts that set the step's
   algorithmic bytes). */
B2G_API int b2g_stats(b2g_ctx* ctx, int64_t* out, int32_t n, int32_t reset);

/* The contact listener of the last step (ContactManager.begin_contact,
   end_contact, pre_solve, post_solve; contact_manager.py:47-98): the records of
   the kinds enabled by B2G_PARAM_EVENTS (layout in b2g_layout.h), in the order
   the reference calls the listener: by world, then phase (ContactManager.collide's
   loop over the contacts, contact.py:262-268; the end_contact of the contacts it
   destroys, contact_manager.py:160-172; Island.report, island.py:245-252).
   *count receives the number of records; with more than cap nothing is written
   and B2G_E_CAPACITY is returned.  b2g_step_n keeps its last step's records. */
B2G_API int b2g_events(b2g_ctx* ctx, int32_t* out_i, double* out_d, int32_t cap, int32_t* count);

/* One step in two parts, so that the contact listener's begin_contact,
   end_contact and pre_solve run between them as in the reference, where
   pre_solve may disable a contact for this step (Contact.enabled = False,
   contact.py:116-127; testbed/one_sided_platform.py): b2g_step_collide runs the
   step up to and with ContactManager.collide (the pending new-fixture pair
   search, the narrow phase; world.py:316-338), b2g_events then holds that
   part's records, b2g_disable_contacts applies the listener's verdicts, and
   b2g_step_solve runs the rest (world.py:340-354: islands, solve, Island.report,
   synchronisation, find_new_contacts).  The two parts give the same state as
   b2g_step.  No other step call is accepted in between. */
B2G_API int b2g_step_collide(b2g_ctx* ctx, double dt, int vel_iters, int pos_iters);
B2G_API int b2g_step_solve(b2g_ctx* ctx);

/* Parity seam of the solver: the island search and Island.solve of World._solve
   (world.py:494-580, island.py:65-179, the ContactSolver of contact.py:375-864 and
   the joints) on the state as it stands, dt > 0: no collide (manifolds as
   uploaded), no fixture synchronisation, no pair search, no time-of-impact pass,
   forces not cleared.  Separates a solver difference from narrowphase or
   pair-search differences. */
B2G_API int b2g_solve_islands(b2g_ctx* ctx, double dt, int vel_iters, int pos_iters);

/* contact.enabled = False for this step: contact slot slot[k] of world world[k]
   (the slot is a phase-0 event's key / 4, its place in the world's contact list
   during collide).  Only between b2g_step_collide and b2g_step_solve. */
B2G_API int b2g_disable_contacts(b2g_ctx* ctx, int32_t n, const int32_t* world, const int32_t* slot);

/* The new-fixture pair search of the next step run ahead of it
   (World.step -> ContactManager.find_new_contacts when fixtures were added,
   world.py:316-319), so that a custom contact filter can see the new contacts
   before the step uses them.  The next b2g_step does not repeat it.  With no
   search pending, nothing is new. */
B2G_API int b2g_pair_search(b2g_ctx* ctx);

/* The contacts created by each world's last pair search (the step's
   ContactManager.find_new_contacts, contact_manager.py:102-158, or
   b2g_pair_search), 4 int32 per contact: world, fixture A, fixture B (world-
   local, as the contact holds them), flags: B2G_CF_SWAPPED when the contact
   registry swapped the broad phase's pair (contact.py:305-326), so that the
   pair the reference hands its contact_filter is (B, A) instead of (A, B);
   B2G_NC_DROPPED once dropped.  Rows by world, then creation order.  *count
   receives the number of rows; with more than cap nothing is written and
   B2G_E_CAPACITY is returned.  A custom contact_filter (world.py:124-136,
   contact_manager.py:144-146) runs on these rows on the host; with
   B2G_WF_CUSTOM_FILTER set the pair search applies no filter of its own. */
B2G_API int b2g_new_contacts(b2g_ctx* ctx, int32_t* out, int32_t cap, int32_t* count);

/* Removes contacts created by the last pair search, before the next step:
   contact index[k] (0-based among world[k]'s rows of b2g_new_contacts) of
   world world[k].  The rest of the contact list keeps its order.  The
   counterpart of the reference's filter returning False (contact_manager.py:
   144-146): that contact is never created.  B2G_E_INVALID for an index outside
   the world's new contacts. */
B2G_API int b2g_drop_contacts(b2g_ctx* ctx, int32_t n, const int32_t* world, const int32_t* index);

/* Parity seam, narrow phase: evaluates n contacts (fixture records A and B
   already oriented as the contact registry orders them, contact.py:305-326,
   at body transforms xf = (px, py, cos, sin)) from zeroed manifolds, like
   Contact.evaluate (contact.py:343-373, collision.py).  Writes one contact
   record per contact (manifold fields only). */
B2G_API int b2g_narrowphase(int precision, int n, const double* fa_d, const int32_t* fa_i, const double* fb_d,
                            const int32_t* fb_i, const double* xfa, const double* xfb, double* con_d,
                            int32_t* con_i);

/* Parity seam, broad phase: the pair set BroadPhase.update_pairs
   (broadphase.py:83-109) reports for fat AABBs fat[n][4] (lx, ly, ux, uy) and
   a move buffer, as (moved proxy, partner) in canonical order (move-buffer
   position, then partner index).  out holds cap pairs (2 * cap ints). */
B2G_API int b2g_pairs(int n, const double* fat, int nmove, const int32_t* move, int32_t* out, int cap,
                      int32_t* count);

/* The pair seam of b2g_pairs without any allocation, for callers that cannot
   allocate device memory (code under HIP graph capture, serving loops, harnesses
   that forbid hipMalloc).  No counterpart in the reference.  Inputs and outputs
   are device pointers the caller owns: fat[n][4], move[nmove], out[2 * cap],
   *count; the work is queued on `stream` (a hipStream_t, or NULL for the null
   stream) and nothing is synchronised.  *count receives the number of pairs
   (only the first cap are written, in canonical order: move-buffer position,
   then partner), or -1 when a move entry is out of range or repeated.  The
   scratch (move positions, per-proxy counts and offsets) lives in a fixed
   per-device array of 8 MiB: n and nmove are limited to 1 << 20 each, larger
   sizes return B2G_E_INVALID before anything is launched.  Calls that overlap
   on one GPU share that array and race. */
B2G_API int b2g_pairs_noalloc(int n, const double* fat, int nmove, const int32_t* move, int32_t* out, int cap,
                              int32_t* count, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* B2G_H */
