/*
 * b2g_layout.h -- record layout of a batch of 2-D rigid-body worlds as it
 * crosses the C-ABI (host arrays, caller-owned).
 *
 * A batch holds W independent worlds.  Every world owns a contiguous range of
 * bodies, fixtures (one broad-phase proxy per fixture) and contacts, given by
 * the offset arrays body_off / fix_off / con_off (length W+1).  All references
 * inside a world (fixture -> body, contact -> fixture, move buffer -> fixture)
 * are world-local indices.
 *
 * Records are fixed-stride rows of float64 and int32.  The field meanings are
 * those of the pypybox2d objects they mirror:
 *   body    : body.py:39-116 (Body), common.py:870-923 (Sweep)
 *   fixture : fixture.py:55-85 (Fixture), shapes.py (Circle/Polygon/Edge),
 *             dynamictree.py:115-134 (the fat AABB kept by the tree node)
 *   contact : contact.py:55-90 (Contact), contact_util.py:50-193 (Manifold)
 *   world   : world.py:64-91 (World), world.py:53-62 (TimeStep)
 * Contacts are stored in ContactManager.contacts order (contact_manager.py:154),
 * which is also the per-body contact-list order used by the island search.
 */
#ifndef B2G_LAYOUT_H
#define B2G_LAYOUT_H

#include <stdint.h>

#define B2G_MAX_POLY_VERTS 8

/* shape type codes: shapes.py:884-889 (Circle=0, Polygon=1, Loop=2, Edge=3) */
#define B2G_SHAPE_CIRCLE 0
#define B2G_SHAPE_POLYGON 1
#define B2G_SHAPE_LOOP 2
#define B2G_SHAPE_EDGE 3

/* body type codes: body.py:36-38 */
#define B2G_STATIC 0
#define B2G_DYNAMIC 1
#define B2G_KINEMATIC 2

/* manifold type codes: contact_util.py:148-150 */
#define B2G_M_CIRCLES 0
#define B2G_M_FACE_A 1
#define B2G_M_FACE_B 2

/* ---- body record: float64 x B2G_BODY_ND ---- */
enum {
  B2G_BD_XF_PX = 0, B2G_BD_XF_PY, B2G_BD_XF_C, B2G_BD_XF_S, /* body._xf: p, R=[[c,-s],[s,c]] */
  B2G_BD_LC_X, B2G_BD_LC_Y,                               /* sweep.local_center */
  B2G_BD_C0_X, B2G_BD_C0_Y, B2G_BD_C_X, B2G_BD_C_Y,       /* sweep.c0, sweep.c */
  B2G_BD_A0, B2G_BD_A,                                     /* sweep.a0, sweep.a */
  B2G_BD_V_X, B2G_BD_V_Y, B2G_BD_W,                        /* linear / angular velocity */
  B2G_BD_F_X, B2G_BD_F_Y, B2G_BD_TORQUE,                   /* accumulated force / torque */
  B2G_BD_MASS, B2G_BD_INV_MASS, B2G_BD_I, B2G_BD_INV_I,
  B2G_BD_GRAV_SCALE, B2G_BD_LIN_DAMP, B2G_BD_ANG_DAMP,
  B2G_BD_SLEEP_TIME,
  B2G_BODY_ND = 28
};
/* ---- body record: int32 x B2G_BODY_NI ---- */
enum { B2G_BI_TYPE = 0, B2G_BI_FLAGS, B2G_BI_ISLAND, B2G_BI_PAD, B2G_BODY_NI = 4 };
#define B2G_BF_AWAKE 1u
#define B2G_BF_ALLOW_SLEEP 2u
#define B2G_BF_FIXED_ROT 4u
#define B2G_BF_ACTIVE 8u
#define B2G_BF_BULLET 16u

/* ---- fixture record: float64 x B2G_FIX_ND ---- */
enum {
  B2G_FD_FRICTION = 0, B2G_FD_RESTITUTION, B2G_FD_DENSITY, B2G_FD_RADIUS,
  B2G_FD_FAT_LX, B2G_FD_FAT_LY, B2G_FD_FAT_UX, B2G_FD_FAT_UY, /* tree-node (fat) AABB */
  B2G_FD_CENTROID_X, B2G_FD_CENTROID_Y,                       /* Polygon._centroid */
  B2G_FD_CIRCLE_PX, B2G_FD_CIRCLE_PY,                         /* Circle.position */
  B2G_FD_EDGE_V0X, B2G_FD_EDGE_V0Y, B2G_FD_EDGE_V1X, B2G_FD_EDGE_V1Y,
  B2G_FD_EDGE_V2X, B2G_FD_EDGE_V2Y, B2G_FD_EDGE_V3X, B2G_FD_EDGE_V3Y,
  B2G_FD_VERTS = 20,                                         /* 8 x (x,y) */
  B2G_FD_NORMALS = B2G_FD_VERTS + 2 * B2G_MAX_POLY_VERTS,    /* 8 x (x,y) */
  B2G_FIX_ND = 56
};
/* ---- fixture record: int32 x B2G_FIX_NI ---- */
enum {
  B2G_FI_BODY = 0, B2G_FI_SHAPE, B2G_FI_NVERTS, B2G_FI_EDGE_FLAGS,
  B2G_FI_CATEGORY, B2G_FI_MASK, B2G_FI_GROUP, B2G_FI_SENSOR,
  B2G_FIX_NI = 8
};
#define B2G_EF_HAS_V0 1u
#define B2G_EF_HAS_V3 2u

/* ---- contact record: float64 x B2G_CON_ND ---- */
enum {
  B2G_CD_FRICTION = 0, B2G_CD_RESTITUTION,
  B2G_CD_LN_X, B2G_CD_LN_Y, B2G_CD_LP_X, B2G_CD_LP_Y,       /* manifold local_normal / local_point */
  B2G_CD_P0_X, B2G_CD_P0_Y, B2G_CD_P0_NI, B2G_CD_P0_TI,     /* points[0] */
  B2G_CD_P1_X, B2G_CD_P1_Y, B2G_CD_P1_NI, B2G_CD_P1_TI,     /* points[1] */
  B2G_CON_ND = 16
};
/* ---- contact record: int32 x B2G_CON_NI ---- */
enum {
  B2G_CI_FA = 0, B2G_CI_FB, B2G_CI_FLAGS, B2G_CI_MTYPE, B2G_CI_PC,
  B2G_CI_ID0, B2G_CI_ID1, B2G_CI_PAD,
  B2G_CON_NI = 8
};
#define B2G_CF_TOUCHING 1u
#define B2G_CF_ENABLED 2u
#define B2G_CF_FILTER 4u
/* b2g_new_contacts only: the contact registry swapped the broad phase's pair when the
   contact was created (contact.py:305-326): the pair was reported as (fixture B, fixture A) */
#define B2G_CF_SWAPPED 8u
/* with B2G_WF_CUSTOM_FILTER, on a contact flagged B2G_CF_FILTER: the host ran the
   custom filter for it (ContactManager.collide, contact_manager.py:200-205) and
   it returned False; collide destroys the contact */
#define B2G_CF_REJECT 16u
/* b2g_new_contacts: the contact has been dropped (b2g_drop_contacts) */
#define B2G_NC_DROPPED 16
/* a contact-point id (contact_util.py:78-79) packed as 4 bytes:
   id[0] | id[1] << 8 | id[2] << 16 | id[3] << 24 */

/* ---- world record ---- */
enum { B2G_WD_GX = 0, B2G_WD_GY, B2G_WD_INV_DT0, B2G_WD_PAD, B2G_WORLD_ND = 4 };
/* B2G_WI_NEW_CONTACTS (output only): contacts created by the world's last pair search
   (find_new_contacts); they are the last ones of its contact list */
enum { B2G_WI_FLAGS = 0, B2G_WI_NEW_CONTACTS, B2G_WI_PAD2, B2G_WI_PAD3, B2G_WORLD_NI = 4 };
#define B2G_WF_ALLOW_SLEEP 1u
#define B2G_WF_WARM_STARTING 2u
#define B2G_WF_NEW_FIXTURE 4u
#define B2G_WF_KEEP_FORCES 8u     /* auto_clear_forces off */
/* a custom contact_filter replaces World.default_contact_filter (world.py:86-89): the pair
   search creates every pair that passes Body._should_collide, and the host runs its filter on
   the new contacts and drops the rejected ones (b2g_drop_contacts) before the next step, and
   decides the contacts flagged for filtering (B2G_CF_REJECT) */
#define B2G_WF_CUSTOM_FILTER 16u
/* continuous physics on (World.continuous_physics, world.py:345-346): after each step with
   dt > 0, World.solve_toi (world.py:594-821) sub-steps the first impacts of fast bodies */
#define B2G_WF_CONTINUOUS 32u

/* ---- contact event record (b2g_events): int32 x B2G_EV_NI, float64 x B2G_EV_ND ----
   One callback of the contact listener (contact.py:210-268, contact_manager.py:160-172,
   island.py:245-252).  kind: B2G_EV_BEGIN / END / PRE_SOLVE / POST_SOLVE.  Events of one
   step are ordered by (world, phase, key): phase 0, key 4 x contact + 0/1/2 for begin, end,
   pre_solve in ContactManager.collide's loop; phase 1, key contact for the end_contact of a
   contact that collide destroys while touching; phase 2, key = the contact's position in the
   island lists for post_solve.  The manifold fields hold the contact's manifold at the call;
   the old_* fields hold, for pre_solve, the manifold before Contact.update, and for
   post_solve the impulses handed to the listener: old pc of them (one per velocity-constraint
   point), in the old P0/P1 normal and tangent impulse fields. */
#define B2G_EV_BEGIN 0
#define B2G_EV_END 1
#define B2G_EV_PRE_SOLVE 2
#define B2G_EV_POST_SOLVE 3
enum {
  B2G_EI_WORLD = 0, B2G_EI_KIND, B2G_EI_PHASE, B2G_EI_KEY, B2G_EI_FA, B2G_EI_FB, B2G_EI_FLAGS,
  B2G_EI_MTYPE, B2G_EI_PC, B2G_EI_ID0, B2G_EI_ID1,
  B2G_EI_OLD_MTYPE, B2G_EI_OLD_PC, B2G_EI_OLD_ID0, B2G_EI_OLD_ID1, B2G_EI_PAD,
  B2G_EV_NI = 16
};
enum {
  B2G_ED_LN_X = 0, B2G_ED_LN_Y, B2G_ED_LP_X, B2G_ED_LP_Y,
  B2G_ED_P0_X, B2G_ED_P0_Y, B2G_ED_P0_NI, B2G_ED_P0_TI,
  B2G_ED_P1_X, B2G_ED_P1_Y, B2G_ED_P1_NI, B2G_ED_P1_TI,
  B2G_ED_OLD = 12, /* the old manifold, the same 12 fields */
  B2G_EV_ND = 24
};

/* A batch description.  Every pointer is a caller-owned host array; sizes are
   given by the offset arrays (each of length n_worlds + 1). */
typedef struct b2g_batch {
  int32_t n_worlds;
  const int32_t *body_off, *fix_off, *con_off, *move_off;
  const double *world_d;  /* n_worlds x B2G_WORLD_ND */
  const int32_t *world_i; /* n_worlds x B2G_WORLD_NI */
  const double *body_d;   /* n_bodies x B2G_BODY_ND */
  const int32_t *body_i;  /* n_bodies x B2G_BODY_NI */
  const double *fix_d;    /* n_fixtures x B2G_FIX_ND */
  const int32_t *fix_i;   /* n_fixtures x B2G_FIX_NI */
  const double *con_d;    /* n_contacts x B2G_CON_ND */
  const int32_t *con_i;   /* n_contacts x B2G_CON_NI */
  const int32_t *move;    /* broad-phase move buffer, world-local fixture indices */
} b2g_batch;

/* Output view of a batch: the same arrays, writable.  Body/fixture counts do
   not change during stepping; contact and move-buffer counts do, so their
   arrays carry a capacity (rows) and the offsets are written back. */
typedef struct b2g_batch_out {
  int32_t n_worlds;
  int32_t con_cap, move_cap;
  int32_t *con_off, *move_off; /* written: n_worlds + 1 */
  double *world_d;
  int32_t *world_i;
  double *body_d;
  int32_t *body_i;
  double *fix_d;
  int32_t *fix_i;
  double *con_d;
  int32_t *con_i;
  int32_t *move;
} b2g_batch_out;

#endif /* B2G_LAYOUT_H */
