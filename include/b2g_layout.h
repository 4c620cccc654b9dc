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

/* ---- joint records: int32 x B2G_JOINT_NI, float64 x B2G_JOINT_ND ----
   One joint of World.joints (world.py:252-283), in the world's joint-list order, which is
   also the order of each body's joint list (Body.joints: add_joint appends to body_a's list,
   then body_b's).  The double record holds the joint object's numbers: its parameters,
   its accumulated impulses and the values its solver computes each step (the reference
   keeps them as attributes; some are read in later steps or by the reaction getters). */
#define B2G_JT_REVOLUTE 1  /* joints/revolute.py */
#define B2G_JT_PRISMATIC 2 /* joints/prismatic.py */
#define B2G_JT_DISTANCE 3  /* joints/distance.py */
#define B2G_JT_PULLEY 4    /* joints/pulley.py */
#define B2G_JT_MOUSE 5     /* joints/mouse.py: body_a is none (-1) */
#define B2G_JT_GEAR 6      /* joints/gear.py: bodies a, b, c, d as the constructor sets them */
#define B2G_JT_WHEEL 7     /* joints/wheel.py */
#define B2G_JT_WELD 8      /* joints/weld.py */
#define B2G_JT_FRICTION 9  /* joints/friction.py */
#define B2G_JT_ROPE 10     /* joints/rope.py */
enum {
  B2G_JI_TYPE = 0, B2G_JI_BODY_A, B2G_JI_BODY_B, /* world-local body indices, -1 for none */
  B2G_JI_FLAGS, B2G_JI_STATE,                    /* limit state (joint.py:35), rope _state */
  B2G_JI_BODY_C, B2G_JI_BODY_D,                  /* gear joint: bodies c and d */
  B2G_JI_PAD,
  B2G_JOINT_NI = 8
};
#define B2G_JF_COLLIDE_CONNECTED 1u
#define B2G_JF_LIMIT 2u      /* limit_enabled */
#define B2G_JF_MOTOR 4u      /* motor_enabled */
#define B2G_JF_REVOLUTE_A 8u /* gear: joint_a is a revolute joint */
#define B2G_JF_REVOLUTE_B 16u
/* limit states (joint.py:35) */
#define B2G_LIMIT_INACTIVE 0
#define B2G_LIMIT_AT_LOWER 1
#define B2G_LIMIT_AT_UPPER 2
#define B2G_LIMIT_EQUAL 3
enum {
  /* parameters and accumulated impulses */
  B2G_JD_LA_X = 0, B2G_JD_LA_Y, B2G_JD_LB_X, B2G_JD_LB_Y, /* local anchors a, b */
  B2G_JD_REF,                      /* reference angle; gear: reference_angle_a */
  B2G_JD_IMP_X, B2G_JD_IMP_Y, B2G_JD_IMP_Z,
  /* impulse: Vec3 (revolute, prismatic, weld), Vec2 (mouse), scalar in X (distance, pulley,
     gear, wheel, rope); friction: linear impulse in X, Y and angular impulse in Z */
  B2G_JD_MOTOR_IMP,                /* 8: revolute, prismatic, wheel */
  B2G_JD_MOTOR_SPEED,
  B2G_JD_MAX_MOTOR,                /* 10: max motor torque (revolute, wheel), force (prismatic);
                                      max_force (friction, mouse) */
  B2G_JD_LOWER, B2G_JD_UPPER,      /* limits */
  B2G_JD_AXIS_X, B2G_JD_AXIS_Y,    /* 13: local x axis (prismatic, wheel); gear: local_axis_c */
  B2G_JD_PERP_X, B2G_JD_PERP_Y,    /* 15: local y axis (prismatic, wheel); gear: local_axis_d */
  B2G_JD_FREQ, B2G_JD_DAMP,        /* 17: frequency, damping ratio (distance, wheel, mouse) */
  B2G_JD_LENGTH,                   /* 19: distance length, rope max length; pulley, gear constant */
  B2G_JD_RATIO,                    /* 20: pulley, gear */
  B2G_JD_G0_X, B2G_JD_G0_Y, B2G_JD_G1_X, B2G_JD_G1_Y,
  /* 21: pulley ground anchors a, b; mouse target (G0); gear local anchors c, d */
  B2G_JD_SPRING_IMP,               /* 25: wheel */
  B2G_JD_MAX_TORQUE,               /* 26: friction */
  B2G_JD_REF_B,                    /* 27: gear reference_angle_b */
  B2G_JD_LEN_A, B2G_JD_LEN_B,      /* 28: pulley length_a, length_b */
  /* the solver's values of the last step */
  B2G_JD_RA_X = 32, B2G_JD_RA_Y, B2G_JD_RB_X, B2G_JD_RB_Y, /* lever arms _ra, _rb */
  B2G_JD_S = 36,                   /* per-type values: */
  /* revolute: _mass (Mat33 by columns), _motor_mass */
  B2G_JD_REV_MASS = B2G_JD_S, B2G_JD_REV_MOTOR_MASS = B2G_JD_S + 9,
  /* prismatic: _axis, _perp, _a1, _a2, _s1, _s2, _motor_mass, _k (Mat33 by columns) */
  B2G_JD_PRI_AXIS_X = B2G_JD_S, B2G_JD_PRI_AXIS_Y, B2G_JD_PRI_PERP_X, B2G_JD_PRI_PERP_Y,
  B2G_JD_PRI_A1, B2G_JD_PRI_A2, B2G_JD_PRI_S1, B2G_JD_PRI_S2, B2G_JD_PRI_MOTOR_MASS, B2G_JD_PRI_K,
  /* distance: _u, _mass, _gamma, _bias */
  B2G_JD_DIS_U_X = B2G_JD_S, B2G_JD_DIS_U_Y, B2G_JD_DIS_MASS, B2G_JD_DIS_GAMMA, B2G_JD_DIS_BIAS,
  /* pulley: _ua, _ub, _mass */
  B2G_JD_PUL_UA_X = B2G_JD_S, B2G_JD_PUL_UA_Y, B2G_JD_PUL_UB_X, B2G_JD_PUL_UB_Y, B2G_JD_PUL_MASS,
  /* mouse: _mass (Mat22 by columns), _gamma, _beta, _c */
  B2G_JD_MOU_MASS = B2G_JD_S, B2G_JD_MOU_GAMMA = B2G_JD_S + 4, B2G_JD_MOU_BETA, B2G_JD_MOU_C_X,
  B2G_JD_MOU_C_Y,
  /* wheel: _ax, _ay, _s_ax, _s_bx, _s_ay, _s_by, _mass, _motor_mass, _spring_mass, _bias, _gamma */
  B2G_JD_WHE_AX_X = B2G_JD_S, B2G_JD_WHE_AX_Y, B2G_JD_WHE_AY_X, B2G_JD_WHE_AY_Y, B2G_JD_WHE_S_AX,
  B2G_JD_WHE_S_BX, B2G_JD_WHE_S_AY, B2G_JD_WHE_S_BY, B2G_JD_WHE_MASS, B2G_JD_WHE_MOTOR_MASS,
  B2G_JD_WHE_SPRING_MASS, B2G_JD_WHE_BIAS, B2G_JD_WHE_GAMMA,
  /* weld: _mass (Mat33 by columns) */
  B2G_JD_WEL_MASS = B2G_JD_S,
  /* friction: _linear_mass (Mat22 by columns), _angular_mass */
  B2G_JD_FRI_LMASS = B2G_JD_S, B2G_JD_FRI_AMASS = B2G_JD_S + 4,
  /* rope: _u, _length, _mass */
  B2G_JD_ROP_U_X = B2G_JD_S, B2G_JD_ROP_U_Y, B2G_JD_ROP_LENGTH, B2G_JD_ROP_MASS,
  /* gear (uses the lever-arm slots too): _j_vac, _j_vbd, _j_wa, _j_wb, _j_wc, _j_wd, _mass,
     then the body states it captured at its velocity-constraint setup: linear velocities,
     centres and angles of bodies a, b, c, d */
  B2G_JD_GEA_JVAC_X = B2G_JD_RA_X, B2G_JD_GEA_JVAC_Y, B2G_JD_GEA_JVBD_X, B2G_JD_GEA_JVBD_Y,
  B2G_JD_GEA_JWA, B2G_JD_GEA_JWB, B2G_JD_GEA_JWC, B2G_JD_GEA_JWD, B2G_JD_GEA_MASS,
  B2G_JD_GEA_V = B2G_JD_GEA_MASS + 1, /* 4 x (x, y) */
  B2G_JD_GEA_C = B2G_JD_GEA_V + 8,    /* 4 x (x, y) */
  B2G_JD_GEA_A = B2G_JD_GEA_C + 8,    /* 4 */
  B2G_JOINT_ND = 64
};

/* A batch description.  Every pointer is a caller-owned host array; sizes are
   given by the offset arrays (each of length n_worlds + 1).  The joint arrays
   may be NULL (no joints). */
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
  const int32_t *joint_off; /* n_worlds + 1, or NULL */
  const double *joint_d;    /* n_joints x B2G_JOINT_ND */
  const int32_t *joint_i;   /* n_joints x B2G_JOINT_NI */
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
  double *joint_d;  /* n_joints x B2G_JOINT_ND (joints do not change during stepping), or NULL */
  int32_t *joint_i; /* n_joints x B2G_JOINT_NI, or NULL */
} b2g_batch_out;

#endif /* B2G_LAYOUT_H */
