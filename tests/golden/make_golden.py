"""Writes the golden fixtures of the contact pipeline (tests/golden/*.npz).

Run once in the build container:  python tests/golden/make_golden.py [names]
(names: pyramid stacks mixed filtered listeners continuous joints queries narrowphase; default all)

It imports pypybox2d from /root/reference in its default pure-Python flavour
(the reference tree ships no compiled _common) and records, for small scenes:
  * the initial world state right after construction, in the batch record
    layout of include/b2g_layout.h;
  * for every step, the contacts that step created, in the order the
    reference appended them (its pair set iterates by object address,
    broadphase.py:98-109, so this order is recorded, not recomputed);
  * full world states after selected steps.
and, for the narrow phase, random shape pairs with the manifold each collider
returns.  Nothing here is used at run time on the GPU host; the fixtures are
data.
"""
from __future__ import annotations

import math
import os
import sys
from copy import copy

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
sys.path.insert(0, "/root/reference")

import pypybox2d as b2  # noqa: E402
from pypybox2d import collision as b2col  # noqa: E402
from pypybox2d import contact_util as b2cu  # noqa: E402

from pypybox2d_amd import layout as L  # noqa: E402
from pypybox2d_amd import scenes  # noqa: E402


# --------------------------------------------------------------------------
# scene spec -> reference world
# --------------------------------------------------------------------------
def ref_shape(spec):
    kind = spec[0]
    if kind == "box":
        return b2.Polygon(box=(spec[1], spec[2]))
    if kind == "circle":
        return b2.Circle(radius=spec[1])
    if kind == "polygon":
        return b2.Polygon(vertices=spec[1])
    if kind == "edge":
        return b2.Edge(spec[1], spec[2])
    raise ValueError(kind)


def build_ref_world(scene, continuous_physics=False):
    world = b2.World(gravity=scene["gravity"], do_sleep=scene["do_sleep"],
                     warm_starting=scene["warm_starting"], continuous_physics=continuous_physics)
    for bs in scene["bodies"]:
        fixtures = [b2.Fixture(ref_shape(f["shape"]), density=f["density"], friction=f["friction"],
                               restitution=f["restitution"], category_bits=scenes.fixture_filter(f)[0],
                               mask_bits=scenes.fixture_filter(f)[1], group_index=scenes.fixture_filter(f)[2])
                    for f in bs["fixtures"]]
        kw = dict(position=bs["position"], angle=bs["angle"], linear_velocity=bs["linear_velocity"],
                  angular_velocity=bs["angular_velocity"], fixtures=fixtures, bullet=bs.get("bullet", False))
        if bs["type"] == "static":
            world.create_static_body(**kw)
        elif bs["type"] == "kinematic":
            world.create_kinematic_body(**kw)
        else:
            world.create_dynamic_body(**kw)
    made = []
    for js in scene.get("joints", ()):
        kw = {k: v for k, v in js.items() if k not in ("type", "bodies", "joints")}
        create = getattr(world, "create_%s_joint" % js["type"])
        if js["type"] == "gear":
            made.append(create(made[js["joints"][0]], made[js["joints"][1]], **kw))
        else:
            made.append(create(*[world.bodies[i] for i in js["bodies"]], **kw))
    return world


# --------------------------------------------------------------------------
# reference joints -> joint records (b2g_layout.h B2G_JI_* / B2G_JD_*)
# --------------------------------------------------------------------------
_JT = {"RevoluteJoint": L.JT_REVOLUTE, "PrismaticJoint": L.JT_PRISMATIC, "DistanceJoint": L.JT_DISTANCE,
       "PulleyJoint": L.JT_PULLEY, "MouseJoint": L.JT_MOUSE, "GearJoint": L.JT_GEAR, "WheelJoint": L.JT_WHEEL,
       "WeldJoint": L.JT_WELD, "FrictionJoint": L.JT_FRICTION, "RopeJoint": L.JT_ROPE}


def joint_record(j, bidx):
    """The joint object's attributes in the record layout (attributes a joint sets only in its
    first velocity-constraint setup read as zero before it)."""
    jd = np.zeros(L.JOINT_ND)
    ji = np.zeros(L.JOINT_NI, np.int32)
    t = _JT[type(j).__name__]
    ji[L.JI_TYPE] = t
    ji[L.JI_BODY_A] = bidx[id(j._body_a)] if j._body_a is not None else -1
    ji[L.JI_BODY_B] = bidx[id(j._body_b)]
    ji[L.JI_FLAGS] = ((L.JF_COLLIDE_CONNECTED if j._collide_connected else 0)
                      | (L.JF_LIMIT if getattr(j, "_limit_enabled", False) else 0)
                      | (L.JF_MOTOR if getattr(j, "_motor_enabled", False) else 0))

    def g(name, default=0.0):
        return getattr(j, name, default)

    def put2(k, v):
        if v is not None:
            jd[k], jd[k + 1] = v[0], v[1]

    def put3(k, v):
        jd[k], jd[k + 1], jd[k + 2] = v.x, v.y, v.z

    def put33(k, m):
        if m is not None and not isinstance(m, float):
            put3(k, m.col1); put3(k + 3, m.col2); put3(k + 6, m.col3)

    def put22(k, m):
        if m is not None and not isinstance(m, float):
            jd[k], jd[k + 1], jd[k + 2], jd[k + 3] = m._col1.x, m._col1.y, m._col2.x, m._col2.y

    if t != L.JT_GEAR:
        put2(L.JD_LA_X, g("_local_anchor_a", None))
        put2(L.JD_LB_X, g("_local_anchor_b", None))
        put2(L.JD_RA_X, g("_ra", None))
        put2(L.JD_RB_X, g("_rb", None))
    if t in (L.JT_REVOLUTE, L.JT_PRISMATIC, L.JT_WELD):
        jd[L.JD_REF] = j._reference_angle
        put3(L.JD_IMP_X, j._impulse)
    if t in (L.JT_REVOLUTE, L.JT_PRISMATIC):
        ji[L.JI_STATE] = j._limit_state
        jd[L.JD_MOTOR_IMP], jd[L.JD_MOTOR_SPEED] = j._motor_impulse, j._motor_speed
    if t == L.JT_REVOLUTE:
        jd[L.JD_MAX_MOTOR] = j._max_motor_torque
        jd[L.JD_LOWER], jd[L.JD_UPPER] = j._lower_angle, j._upper_angle
        put33(L.JD_REV_MASS, j._mass)
        jd[L.JD_REV_MOTOR_MASS] = g("_motor_mass")
    elif t == L.JT_PRISMATIC:
        jd[L.JD_MAX_MOTOR] = j._max_motor_force
        jd[L.JD_LOWER], jd[L.JD_UPPER] = j._lower_limit, j._upper_limit
        put2(L.JD_AXIS_X, j._local_x_axis)
        put2(L.JD_PERP_X, j._local_y_axis)
        put2(L.JD_PRI_AXIS_X, j._axis)
        put2(L.JD_PRI_PERP_X, j._perp)
        jd[L.JD_PRI_A1], jd[L.JD_PRI_A2], jd[L.JD_PRI_S1], jd[L.JD_PRI_S2] = j._a1, j._a2, j._s1, j._s2
        jd[L.JD_PRI_MOTOR_MASS] = j._motor_mass
        put33(L.JD_PRI_K, j._k)
    elif t == L.JT_DISTANCE:
        jd[L.JD_LENGTH], jd[L.JD_FREQ], jd[L.JD_DAMP] = j._length, j._frequency, j._damping_ratio
        jd[L.JD_IMP_X] = j._impulse
        put2(L.JD_DIS_U_X, j._u)
        jd[L.JD_DIS_MASS], jd[L.JD_DIS_GAMMA], jd[L.JD_DIS_BIAS] = j._mass, j._gamma, j._bias
    elif t == L.JT_PULLEY:
        put2(L.JD_G0_X, j._ground_anchor_a)
        put2(L.JD_G1_X, j._ground_anchor_b)
        jd[L.JD_LEN_A], jd[L.JD_LEN_B] = j._length_a, j._length_b
        jd[L.JD_LENGTH], jd[L.JD_RATIO] = j._constant, j._ratio
        jd[L.JD_IMP_X] = j._impulse
        put2(L.JD_PUL_UA_X, g("_ua", None))
        put2(L.JD_PUL_UB_X, g("_ub", None))
        jd[L.JD_PUL_MASS] = j._mass
    elif t == L.JT_MOUSE:
        put2(L.JD_G0_X, j._target)
        jd[L.JD_MAX_MOTOR], jd[L.JD_FREQ], jd[L.JD_DAMP] = j._max_force, j._frequency, j._damping_ratio
        put2(L.JD_IMP_X, j._impulse)
        put22(L.JD_MOU_MASS, g("_mass", None))
        jd[L.JD_MOU_GAMMA], jd[L.JD_MOU_BETA] = j._gamma, j._beta
        put2(L.JD_MOU_C_X, g("_c", None))
    elif t == L.JT_GEAR:
        ji[L.JI_BODY_C] = bidx[id(j._body_c)]
        ji[L.JI_BODY_D] = bidx[id(j._body_d)]
        ji[L.JI_FLAGS] |= (L.JF_REVOLUTE_A if j._revolute_a else 0) | (L.JF_REVOLUTE_B if j._revolute_b else 0)
        put2(L.JD_LA_X, j._local_anchor_a)
        put2(L.JD_LB_X, j._local_anchor_b)
        put2(L.JD_G0_X, j._local_anchor_c)
        put2(L.JD_G1_X, j._local_anchor_d)
        put2(L.JD_AXIS_X, j._local_axis_c)
        put2(L.JD_PERP_X, j._local_axis_d)
        jd[L.JD_REF], jd[L.JD_REF_B] = j._reference_angle_a, j._reference_angle_b
        jd[L.JD_RATIO], jd[L.JD_LENGTH], jd[L.JD_IMP_X] = j._ratio, j._constant, j._impulse
        if hasattr(j, "_j_vac"):
            put2(L.JD_GEA_JVAC_X, j._j_vac)
            put2(L.JD_GEA_JVBD_X, j._j_vbd)
            jd[L.JD_GEA_JWA], jd[L.JD_GEA_JWB], jd[L.JD_GEA_JWC], jd[L.JD_GEA_JWD] = j._j_wa, j._j_wb, j._j_wc, j._j_wd
            jd[L.JD_GEA_MASS] = j._mass
            for k in range(4):
                put2(L.JD_GEA_V + 2 * k, j._lin_vels[k])
                put2(L.JD_GEA_C + 2 * k, j._centers[k])
                jd[L.JD_GEA_A + k] = j._angles[k]
    elif t == L.JT_WHEEL:
        put2(L.JD_AXIS_X, j._local_x_axis)
        put2(L.JD_PERP_X, j._local_y_axis)
        jd[L.JD_FREQ], jd[L.JD_DAMP] = j._frequency, j._damping_ratio
        jd[L.JD_MAX_MOTOR], jd[L.JD_MOTOR_SPEED] = j._max_motor_torque, j._motor_speed
        jd[L.JD_IMP_X], jd[L.JD_MOTOR_IMP], jd[L.JD_SPRING_IMP] = j._impulse, j._motor_impulse, j._spring_impulse
        put2(L.JD_WHE_AX_X, j._ax)
        put2(L.JD_WHE_AY_X, j._ay)
        jd[L.JD_WHE_S_AX], jd[L.JD_WHE_S_BX], jd[L.JD_WHE_S_AY], jd[L.JD_WHE_S_BY] = j._s_ax, j._s_bx, j._s_ay, j._s_by
        jd[L.JD_WHE_MASS], jd[L.JD_WHE_MOTOR_MASS], jd[L.JD_WHE_SPRING_MASS] = j._mass, j._motor_mass, j._spring_mass
        jd[L.JD_WHE_BIAS], jd[L.JD_WHE_GAMMA] = j._bias, j._gamma
    elif t == L.JT_WELD:
        put33(L.JD_WEL_MASS, j._mass)
    elif t == L.JT_FRICTION:
        put2(L.JD_IMP_X, j._linear_impulse)
        jd[L.JD_IMP_Z] = j._angular_impulse
        jd[L.JD_MAX_MOTOR], jd[L.JD_MAX_TORQUE] = j._max_force, j._max_torque
        put22(L.JD_FRI_LMASS, g("_linear_mass", None))
        jd[L.JD_FRI_AMASS] = g("_angular_mass")
    elif t == L.JT_ROPE:
        ji[L.JI_STATE] = j._state
        jd[L.JD_LENGTH], jd[L.JD_IMP_X] = j._max_length, j._impulse
        put2(L.JD_ROP_U_X, j._u)
        jd[L.JD_ROP_LENGTH], jd[L.JD_ROP_MASS] = j._length, j._mass
    return jd, ji


# --------------------------------------------------------------------------
# reference objects -> batch records
# --------------------------------------------------------------------------
def fixture_record(fx):
    fd = np.zeros(L.FIX_ND)
    fi = np.zeros(L.FIX_NI, np.int32)
    sh = fx._shape
    fd[L.FD_FRICTION] = fx._friction
    fd[L.FD_RESTITUTION] = fx._restitution
    fd[L.FD_DENSITY] = fx._density
    fd[L.FD_RADIUS] = sh.radius
    fi[L.FI_SHAPE] = sh._type
    fi[L.FI_CATEGORY] = fx._category_bits
    fi[L.FI_MASK] = fx._mask_bits
    fi[L.FI_GROUP] = fx._group_index
    fi[L.FI_SENSOR] = int(bool(fx._sensor))
    if isinstance(sh, b2.Circle):
        fd[L.FD_CIRCLE_PX], fd[L.FD_CIRCLE_PY] = sh.position.x, sh.position.y
    elif isinstance(sh, b2.Polygon):
        n = len(sh._vertices)
        assert n <= L.MAX_POLY_VERTS
        fi[L.FI_NVERTS] = n
        for k, (v, nm) in enumerate(zip(sh._vertices, sh._normals)):
            fd[L.FD_VERTS + 2 * k], fd[L.FD_VERTS + 2 * k + 1] = v.x, v.y
            fd[L.FD_NORMALS + 2 * k], fd[L.FD_NORMALS + 2 * k + 1] = nm.x, nm.y
        fd[L.FD_CENTROID_X], fd[L.FD_CENTROID_Y] = sh._centroid.x, sh._centroid.y
    elif isinstance(sh, b2.Edge):
        fd[L.FD_EDGE_V1X], fd[L.FD_EDGE_V1Y] = sh._vertex1.x, sh._vertex1.y
        fd[L.FD_EDGE_V2X], fd[L.FD_EDGE_V2Y] = sh._vertex2.x, sh._vertex2.y
        flags = 0
        if sh._vertex0 is not None:
            fd[L.FD_EDGE_V0X], fd[L.FD_EDGE_V0Y] = sh._vertex0.x, sh._vertex0.y
            flags |= L.EF_HAS_V0
        if sh._vertex3 is not None:
            fd[L.FD_EDGE_V3X], fd[L.FD_EDGE_V3Y] = sh._vertex3.x, sh._vertex3.y
            flags |= L.EF_HAS_V3
        fi[L.FI_EDGE_FLAGS] = flags
    else:
        raise ValueError(type(sh))
    if fx._proxies and fx._proxies[0].proxy_ref is not None:
        node = fx._proxies[0].proxy_ref
        fd[L.FD_FAT_LX], fd[L.FD_FAT_LY] = node.aabb._lower_bound.x, node.aabb._lower_bound.y
        fd[L.FD_FAT_UX], fd[L.FD_FAT_UY] = node.aabb._upper_bound.x, node.aabb._upper_bound.y
    return fd, fi


def manifold_record(m, cd, ci):
    ci[L.CI_MTYPE] = m.type
    ci[L.CI_PC] = m.point_count
    cd[L.CD_LN_X], cd[L.CD_LN_Y] = m.local_normal.x, m.local_normal.y
    cd[L.CD_LP_X], cd[L.CD_LP_Y] = m.local_point.x, m.local_point.y
    p0, p1 = m.points[0], m.points[1]
    cd[L.CD_P0_X], cd[L.CD_P0_Y] = p0.local_point.x, p0.local_point.y
    cd[L.CD_P0_NI], cd[L.CD_P0_TI] = p0.normal_impulse, p0.tangent_impulse
    cd[L.CD_P1_X], cd[L.CD_P1_Y] = p1.local_point.x, p1.local_point.y
    cd[L.CD_P1_NI], cd[L.CD_P1_TI] = p1.normal_impulse, p1.tangent_impulse
    ci[L.CI_ID0] = L.pack_id(*p0.id)
    ci[L.CI_ID1] = L.pack_id(*p1.id)


def snapshot(world):
    bodies = world.bodies
    fixtures = [f for b in bodies for f in b.fixtures]
    fidx = {id(f): i for i, f in enumerate(fixtures)}
    bidx = {id(b): i for i, b in enumerate(bodies)}
    nb, nf = len(bodies), len(fixtures)
    bd = np.zeros((nb, L.BODY_ND))
    bi = np.zeros((nb, L.BODY_NI), np.int32)
    for i, b in enumerate(bodies):
        xf, sw = b._xf, b._sweep
        R = xf._rotation
        assert R._col2.x == -R._col1.y and R._col2.y == R._col1.x
        bd[i, L.BD_XF_PX], bd[i, L.BD_XF_PY] = xf._position.x, xf._position.y
        bd[i, L.BD_XF_C], bd[i, L.BD_XF_S] = R._col1.x, R._col1.y
        bd[i, L.BD_LC_X], bd[i, L.BD_LC_Y] = sw.local_center.x, sw.local_center.y
        bd[i, L.BD_C0_X], bd[i, L.BD_C0_Y] = sw.c0.x, sw.c0.y
        bd[i, L.BD_C_X], bd[i, L.BD_C_Y] = sw.c.x, sw.c.y
        bd[i, L.BD_A0], bd[i, L.BD_A] = sw.a0, sw.a
        bd[i, L.BD_V_X], bd[i, L.BD_V_Y] = b._linear_velocity.x, b._linear_velocity.y
        bd[i, L.BD_W] = b._angular_velocity
        bd[i, L.BD_F_X], bd[i, L.BD_F_Y] = b._force.x, b._force.y
        bd[i, L.BD_TORQUE] = b._torque
        bd[i, L.BD_MASS], bd[i, L.BD_INV_MASS] = b._mass, b._inv_mass
        bd[i, L.BD_I], bd[i, L.BD_INV_I] = b._I, b._invI
        bd[i, L.BD_GRAV_SCALE] = b._gravity_scale
        bd[i, L.BD_LIN_DAMP], bd[i, L.BD_ANG_DAMP] = b._linear_damping, b._angular_damping
        bd[i, L.BD_SLEEP_TIME] = b._sleep_time
        bi[i, L.BI_TYPE] = b._type
        bi[i, L.BI_FLAGS] = ((L.BF_AWAKE if b._awake else 0) | (L.BF_ALLOW_SLEEP if b._allow_sleep else 0)
                             | (L.BF_FIXED_ROT if b._fixed_rotation else 0) | (L.BF_ACTIVE if b._active else 0)
                             | (L.BF_BULLET if b._bullet else 0))
        bi[i, L.BI_ISLAND] = -1
    fd = np.zeros((nf, L.FIX_ND))
    fi = np.zeros((nf, L.FIX_NI), np.int32)
    for k, f in enumerate(fixtures):
        fd[k], fi[k] = fixture_record(f)
        fi[k, L.FI_BODY] = bidx[id(f._body)]
    contacts = world.contact_manager.contacts
    nc = len(contacts)
    cd = np.zeros((nc, L.CON_ND))
    ci = np.zeros((nc, L.CON_NI), np.int32)
    for k, c in enumerate(contacts):
        ci[k, L.CI_FA] = fidx[id(c._fixture_a)]
        ci[k, L.CI_FB] = fidx[id(c._fixture_b)]
        ci[k, L.CI_FLAGS] = ((L.CF_TOUCHING if c._touching_flag else 0) | (L.CF_ENABLED if c._enabled_flag else 0)
                             | (L.CF_FILTER if c._filter_flag else 0))
        cd[k, L.CD_FRICTION] = c._friction
        cd[k, L.CD_RESTITUTION] = c._restitution
        manifold_record(c._manifold, cd[k], ci[k])
    move = [fidx[id(p.user_data.fixture)] for p in world.contact_manager.broadphase._move_buffer]
    wd = np.zeros((1, L.WORLD_ND))
    wd[0, L.WD_GX], wd[0, L.WD_GY] = world._gravity.x, world._gravity.y
    wd[0, L.WD_INV_DT0] = world._inv_dt0
    wi = np.zeros((1, L.WORLD_NI), np.int32)
    wi[0, L.WI_FLAGS] = ((L.WF_ALLOW_SLEEP if world._allow_sleep else 0)
                         | (L.WF_WARM_STARTING if world._warm_starting else 0)
                         | (L.WF_NEW_FIXTURE if world._new_fixture else 0)
                         | (0 if world._auto_clear_forces else L.WF_KEEP_FORCES)
                         | (L.WF_CONTINUOUS if world._continuous_physics else 0))
    nj = len(world.joints)
    jd = np.zeros((nj, L.JOINT_ND))
    ji = np.zeros((nj, L.JOINT_NI), np.int32)
    for k, j in enumerate(world.joints):
        jd[k], ji[k] = joint_record(j, bidx)
    return L.BatchArrays(np.array([0, nb], np.int32), np.array([0, nf], np.int32),
                         np.array([0, nc], np.int32), np.array([0, len(move)], np.int32),
                         wd, wi, bd, bi, fd, fi, cd, ci, np.array(move, np.int32),
                         np.array([0, nj], np.int32), jd, ji), fidx


_toi_events = [0]


def _count_toi_events():
    """Counts Island.solve_toi calls (one per TOI event) for the recording; the reference's own
    method runs unchanged."""
    from pypybox2d import island as b2island
    if getattr(b2island.Island.solve_toi, "_counted", False):
        return
    orig = b2island.Island.solve_toi

    def counted(self, *a, **k):
        _toi_events[0] += 1
        return orig(self, *a, **k)
    counted._counted = True
    b2island.Island.solve_toi = counted


def record_run(scene, steps, checkpoints, dt=1.0 / 60, vel_iters=8, pos_iters=3, continuous=False, stats=None):
    """The run's initial state, each step's new contacts in the reference's creation order (hints)
    and the states at the checkpoints.  With continuous physics, `stats` receives per step the
    time_of_impact calls and the TOI events (sub-step island solves)."""
    world = build_ref_world(scene, continuous_physics=continuous)
    init, fidx = snapshot(world)
    hints = []
    states = {}
    if continuous:
        from pypybox2d import toi as b2toi
        _count_toi_events()
    for s in range(1, steps + 1):
        prev = list(world.contact_manager.contacts)  # hold refs: no address reuse
        prev_ids = set(id(c) for c in prev)
        if continuous:
            calls0, events0 = b2toi.toi_calls, _toi_events[0]
        world.step(dt, vel_iters, pos_iters)
        if continuous and stats is not None:
            stats.append((b2toi.toi_calls - calls0, _toi_events[0] - events0))
        new = [c for c in world.contact_manager.contacts if id(c) not in prev_ids]
        hints.append([(fidx[id(c._fixture_a)], fidx[id(c._fixture_b)]) for c in new])
        if s in checkpoints:
            states[s] = snapshot(world)[0]
        del prev
    return init, hints, states


def save_run(path, init, hints, states, meta):
    out = {}
    for name in (L.ALL_FIELDS if init.n_joints else L._FIELDS):
        out["init_" + name] = getattr(init, name)
    off = [0]
    flat = []
    for h in hints:
        flat.extend(h)
        off.append(len(flat))
    out["hint_off"] = np.array(off, np.int32)
    out["hint_pairs"] = np.array(flat, np.int32).reshape(-1, 2)
    steps = sorted(states)
    out["checkpoints"] = np.array(steps, np.int32)
    for s in steps:
        st = states[s]
        for name in ("body_d", "body_i", "fix_d", "con_d", "con_i", "move"):
            out["s%d_%s" % (s, name)] = getattr(st, name)
        if init.n_joints:
            out["s%d_joint_d" % s], out["s%d_joint_i" % s] = st.joint_d, st.joint_i
    for k, v in meta.items():
        out["meta_" + k] = np.array(v)
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


# --------------------------------------------------------------------------
# contact listener and custom contact filter
# --------------------------------------------------------------------------
def _ev_manifold(m, o, oi):
    o[0], o[1], o[2], o[3] = m.local_normal.x, m.local_normal.y, m.local_point.x, m.local_point.y
    p0, p1 = m.points[0], m.points[1]
    o[4], o[5], o[6], o[7] = p0.local_point.x, p0.local_point.y, p0.normal_impulse, p0.tangent_impulse
    o[8], o[9], o[10], o[11] = p1.local_point.x, p1.local_point.y, p1.normal_impulse, p1.tangent_impulse
    oi[0], oi[1], oi[2], oi[3] = m.type, m.point_count, L.pack_id(*p0.id), L.pack_id(*p1.id)


def record_listener_run(scene, steps, checkpoints):
    """A reference run with every contact listener set (contact_manager.begin_contact, end_contact,
    pre_solve, post_solve) and a custom contact_filter (scenes.listener_filter_verdict on the
    fixtures' indices, kept in user_data, and their filter bits).  pre_solve disables contacts
    by scenes.listener_disable_rule.  Each call is recorded in the event layout of b2g_layout.h (the
    phase and key fields, which order the device's records, are -1 here: the list is in call
    order).  After steps 60, 110 and 150 the filter bits of fixtures change (Fixture.refilter):
    a fixture stops colliding, collides again, and two touching fixtures join a negative group;
    the state right after each change is recorded too."""
    world = build_ref_world(scene)
    fixtures = [f for b in world.bodies for f in b.fixtures]
    for k, f in enumerate(fixtures):
        f.user_data = k

    def bits(f):
        return (f._category_bits, f._mask_bits, f._group_index)

    world.contact_filter = lambda fa, fb: scenes.listener_filter_verdict(fa.user_data, bits(fa), fb.user_data, bits(fb))
    cur = {"step": 0}
    rows = []

    def ev(kind, c, old=None, impulses=None):
        ei = np.zeros(L.EV_NI, np.int32)
        ed = np.zeros(L.EV_ND)
        ei[L.EI_KIND], ei[L.EI_PHASE], ei[L.EI_KEY] = kind, -1, -1
        ei[L.EI_FA], ei[L.EI_FB] = c._fixture_a.user_data, c._fixture_b.user_data
        ei[L.EI_FLAGS] = (L.CF_TOUCHING if c._touching_flag else 0) | (L.CF_ENABLED if c._enabled_flag else 0)
        _ev_manifold(c._manifold, ed[:12], ei[L.EI_MTYPE:L.EI_MTYPE + 4])
        if old is not None:
            _ev_manifold(old, ed[L.ED_OLD:], ei[L.EI_OLD_MTYPE:L.EI_OLD_MTYPE + 4])
        if impulses is not None:
            ei[L.EI_OLD_PC] = len(impulses)
            for j, (n, t) in enumerate(impulses):
                ed[L.ED_OLD + L.ED_P0_NI + 4 * j], ed[L.ED_OLD + L.ED_P0_TI + 4 * j] = n, t
        rows.append((cur["step"], ei, ed))

    def pre_solve(c, old):
        ev(L.EV_PRE_SOLVE, c, old=old)
        if scenes.listener_disable_rule(cur["step"], c._fixture_a.user_data, c._fixture_b.user_data):
            c.enabled = False

    cm = world.contact_manager
    cm.begin_contact = lambda c: ev(L.EV_BEGIN, c)
    cm.end_contact = lambda c: ev(L.EV_END, c)
    cm.pre_solve = pre_solve
    cm.post_solve = lambda c, imp: ev(L.EV_POST_SOLVE, c, impulses=imp)
    init, fidx = snapshot(world)
    hints, states, edit_states, edits = [], {}, {}, []

    def touching_partners(f):
        return [c for c in cm.contacts if c._touching_flag and f in (c._fixture_a, c._fixture_b)]

    picked = {}
    for s in range(1, steps + 1):
        cur["step"] = s
        prev = list(cm.contacts)
        prev_ids = set(id(c) for c in prev)
        world.step(1.0 / 60, 8, 3)
        new = [c for c in cm.contacts if id(c) not in prev_ids]
        hints.append([(fidx[id(c._fixture_a)], fidx[id(c._fixture_b)]) for c in new])
        del prev
        if s in checkpoints:
            states[s] = snapshot(world)[0]
        changed = []
        if s == 60:  # the dynamic fixture touching most others stops colliding
            dyn = [f for f in fixtures if f._body._type == b2.Body.DYNAMIC]
            f = max(dyn, key=lambda f: (len(touching_partners(f)), -f.user_data))
            picked["x"], picked["mask"] = f, f._mask_bits
            f.mask_bits = 0
            changed.append(f)
        elif s == 110:  # ... and collides again
            f = picked["x"]
            f.mask_bits = picked["mask"]
            changed.append(f)
        elif s == 150:  # two touching dynamic fixtures join group -3 (never collide)
            pairs = [c for c in cm.contacts if c._touching_flag and c._fixture_a._body._type == b2.Body.DYNAMIC
                     and c._fixture_b._body._type == b2.Body.DYNAMIC]
            c = min(pairs, key=lambda c: (c._fixture_a.user_data, c._fixture_b.user_data))
            for f in (c._fixture_a, c._fixture_b):
                f.group_index = -3
                changed.append(f)
        if changed:
            edit_states[s] = snapshot(world)[0]
            edits.extend((s, f.user_data, f._category_bits, f._mask_bits, f._group_index) for f in changed)
    off = np.zeros(steps + 1, np.int32)
    for st, _, _ in rows:
        off[st] += 1
    off = np.concatenate([[0], np.cumsum(off[1:])]).astype(np.int32)
    return (init, hints, states, edit_states, np.array(edits, np.int32), off,
            np.array([r[1] for r in rows], np.int32).reshape(-1, L.EV_NI),
            np.array([r[2] for r in rows]).reshape(-1, L.EV_ND))


# --------------------------------------------------------------------------
# continuous physics: what World.step does with it on
# --------------------------------------------------------------------------
def record_continuous_run(scene, steps):
    """World.step with continuous_physics on (the reference's default), `steps` times, each call
    caught: the exception it raises (type name and message, or empty), then the world's state,
    _inv_dt0 and _locked; the body states of the same world stepped with continuous physics off;
    and the same for a world without bodies."""
    world = build_ref_world(scene, continuous_physics=True)
    for b in world.bodies:
        b._force = b2.Vec2(1.5, -2.0)  # shows whether the step clears forces
    init, _ = snapshot(world)
    out = {"init_" + n: getattr(init, n) for n in L._FIELDS}
    for s in range(1, steps + 1):
        err = ""
        try:
            world.step(1.0 / 60, 8, 3)
        except Exception as e:  # recorded, not handled
            err = "%s: %s" % (type(e).__name__, e)
        st, _ = snapshot(world)
        for n in ("body_d", "body_i", "fix_d", "con_d", "con_i", "move"):
            out["s%d_%s" % (s, n)] = getattr(st, n)
        out["s%d_error" % s] = np.array(err)
        out["s%d_inv_dt0" % s] = np.array(world._inv_dt0)
        out["s%d_locked" % s] = np.array(bool(world._locked))
    twin = build_ref_world(scene, continuous_physics=False)  # the same world, discrete
    for b in twin.bodies:
        b._force = b2.Vec2(1.5, -2.0)
    for s in range(1, steps + 1):
        twin.step(1.0 / 60, 8, 3)
        out["d%d_body_d" % s] = snapshot(twin)[0].body_d
    empty = b2.World((0, -10), True)
    err = ""
    try:
        empty.step(1.0 / 60, 8, 3)
    except Exception as e:
        err = "%s: %s" % (type(e).__name__, e)
    out["empty_error"] = np.array(err)
    out["empty_inv_dt0"] = np.array(empty._inv_dt0)
    out["meta_steps"] = np.array(steps)
    return out


# --------------------------------------------------------------------------
# narrow phase cases
# --------------------------------------------------------------------------
def _rand_polygon(rng, k=None, rad=None):
    k = int(rng.integers(3, 9)) if k is None else k
    rad = float(rng.uniform(0.3, 1.5)) if rad is None else rad
    ang = np.sort(2.0 * np.pi * np.arange(k) / k + rng.uniform(-0.3, 0.3, k) * (np.pi / k))
    off = rng.uniform(-0.3, 0.3, 2)  # off-centre polygons exercise the zero-centroid start
    return [(float(rad * math.cos(a) + off[0]), float(rad * math.sin(a) + off[1])) for a in ang], rad


def narrowphase_cases(rng, n_per_kind=80):
    cases = []  # (kind, shapeA, shapeB, xfa, xfb)

    def xf_at(p, a):
        return b2.Transform(position=p, angle=a)

    for kind in ("circle_circle", "polygon_circle", "polygon_polygon", "edge_circle", "edge_polygon",
                 "box_box_aligned", "box_edge_aligned"):
        for i in range(n_per_kind):
            aa = float(rng.uniform(-np.pi, np.pi))
            ab = float(rng.uniform(-np.pi, np.pi))
            pa = rng.uniform(-5, 5, 2)
            if kind == "circle_circle":
                ra, rb = rng.uniform(0.2, 1.0, 2)
                A, B = b2.Circle(radius=float(ra), position=tuple(rng.uniform(-.2, .2, 2))), \
                    b2.Circle(radius=float(rb), position=tuple(rng.uniform(-.2, .2, 2)))
                reach = ra + rb
            elif kind == "polygon_circle":
                verts, R = _rand_polygon(rng)
                A = copy(b2.Polygon(vertices=verts))
                rb = float(rng.uniform(0.2, 1.0))
                B = b2.Circle(radius=rb)
                reach = R + rb + 0.3
            elif kind == "polygon_polygon":
                va, Ra = _rand_polygon(rng)
                vb, Rb = _rand_polygon(rng)
                A, B = copy(b2.Polygon(vertices=va)), copy(b2.Polygon(vertices=vb))
                reach = Ra + Rb + 0.3
            elif kind in ("edge_circle", "edge_polygon"):
                L_ = float(rng.uniform(1.0, 8.0))
                v1, v2 = (-L_ / 2, 0.0), (L_ / 2, 0.0)
                v0 = (-L_ / 2 - float(rng.uniform(0.5, 2)), float(rng.uniform(-1, 1))) if rng.random() < 0.4 else None
                v3 = (L_ / 2 + float(rng.uniform(0.5, 2)), float(rng.uniform(-1, 1))) if rng.random() < 0.4 else None
                A = b2.Edge(v1, v2, v0, v3)
                aa = 0.0 if rng.random() < 0.7 else aa
                if kind == "edge_circle":
                    B = b2.Circle(radius=float(rng.uniform(0.2, 1.0)))
                    reach = B.radius + 0.2
                else:
                    vb, Rb = _rand_polygon(rng)
                    B = copy(b2.Polygon(vertices=vb))
                    reach = Rb + 0.2
                # place B near the edge segment
                t = float(rng.uniform(-0.7, 0.7)) * L_
                side = 1.0 if rng.random() < 0.8 else -1.0
                local = np.array([t, side * reach * float(rng.uniform(0.2, 1.1))])
                c, s = math.cos(aa), math.sin(aa)
                pb = pa + np.array([c * local[0] - s * local[1], s * local[0] + c * local[1]])
                cases.append((kind, A, B, xf_at(tuple(pa), aa), xf_at(tuple(pb), ab)))
                continue
            elif kind == "box_box_aligned":
                hx, hy = rng.uniform(0.2, 1.0, 2)
                A = copy(b2.Polygon(box=(float(hx), float(hy))))
                hx2, hy2 = rng.uniform(0.2, 1.0, 2)
                B = copy(b2.Polygon(box=(float(hx2), float(hy2))))
                aa = 0.0
                ab = float(rng.choice([0.0, 0.0, 1e-3, -1e-3, np.pi / 2]))
                gap = float(rng.uniform(-0.02, 0.03))
                dx = float(rng.uniform(-(hx + hx2), hx + hx2))
                pb = pa + np.array([dx, hy + hy2 + gap])
                cases.append((kind, A, B, xf_at(tuple(pa), aa), xf_at(tuple(pb), ab)))
                continue
            else:  # box resting on an axis-aligned edge
                L_ = float(rng.uniform(4.0, 80.0))
                A = b2.Edge((-L_ / 2, 0.0), (L_ / 2, 0.0))
                hx, hy = rng.uniform(0.2, 1.0, 2)
                B = copy(b2.Polygon(box=(float(hx), float(hy))))
                aa = 0.0
                ab = float(rng.choice([0.0, 1e-3, -2e-3, 0.3]))
                gap = float(rng.uniform(-0.02, 0.03))
                pb = pa + np.array([float(rng.uniform(-L_ / 2, L_ / 2)), hy + gap])
                cases.append((kind, A, B, xf_at(tuple(pa), aa), xf_at(tuple(pb), ab)))
                continue
            d = float(rng.uniform(0.3, 1.1)) * reach
            th = float(rng.uniform(-np.pi, np.pi))
            pb = pa + d * np.array([math.cos(th), math.sin(th)])
            cases.append((kind, A, B, xf_at(tuple(pa), aa), xf_at(tuple(pb), ab)))
    return cases


def run_narrowphase(cases):
    fns = {(0, 0): b2col.collide_circles, (1, 1): b2col.collide_polygons, (1, 0): b2col.collide_polygon_circle,
           (3, 0): b2col.collide_edge_circle, (3, 1): b2col.collide_edge_polygon}
    n = len(cases)
    fa_d = np.zeros((n, L.FIX_ND)); fa_i = np.zeros((n, L.FIX_NI), np.int32)
    fb_d = np.zeros((n, L.FIX_ND)); fb_i = np.zeros((n, L.FIX_NI), np.int32)
    xfa = np.zeros((n, 4)); xfb = np.zeros((n, 4))
    cd = np.zeros((n, L.CON_ND)); ci = np.zeros((n, L.CON_NI), np.int32)
    kinds = []
    for k, (kind, A, B, ta, tb) in enumerate(cases):
        fa = b2.Fixture(A)
        fb = b2.Fixture(B)
        fa_d[k], fa_i[k] = fixture_record(fa)
        fb_d[k], fb_i[k] = fixture_record(fb)
        for arr, t in ((xfa, ta), (xfb, tb)):
            arr[k] = (t._position.x, t._position.y, t._rotation._col1.x, t._rotation._col1.y)
        m = b2cu.Manifold()
        fns[(fa._shape._type, fb._shape._type)](m, fa._shape, ta, fb._shape, tb)
        manifold_record(m, cd[k], ci[k])
        kinds.append(kind)
    return dict(fa_d=fa_d, fa_i=fa_i, fb_d=fb_d, fb_i=fb_i, xfa=xfa, xfb=xfb, con_d=cd, con_i=ci,
                kinds=np.array(kinds))


# --------------------------------------------------------------------------
# queries and ray casts
# --------------------------------------------------------------------------
def shape_ray_cases(rng, n_per_kind=120):
    """Random shapes at random transforms and random segments near them, through the reference's
    Shape.ray_cast(p1, p2, 1.0, xf, 0) (shapes.py:147-174, 265-307, 655-702)."""
    rows = []
    for kind in ("circle", "polygon", "edge"):
        for i in range(n_per_kind):
            if kind == "circle":
                sh = b2.Circle(radius=float(rng.uniform(0.2, 1.5)), position=tuple(rng.uniform(-0.5, 0.5, 2)))
            elif kind == "polygon":
                verts, _ = _rand_polygon(rng)
                sh = copy(b2.Polygon(vertices=verts))
            else:
                h = float(rng.uniform(0.5, 4.0))
                sh = b2.Edge((-h, float(rng.uniform(-0.5, 0.5))), (h, float(rng.uniform(-0.5, 0.5))))
            pos = tuple(rng.uniform(-5, 5, 2))
            xf = b2.Transform(position=pos, angle=float(rng.uniform(-np.pi, np.pi)))
            th = float(rng.uniform(-np.pi, np.pi))
            p2 = np.array(pos) + rng.uniform(-1.0, 1.0, 2)
            p1 = p2 + float(rng.uniform(1.0, 6.0)) * np.array([math.cos(th), math.sin(th)])
            if i % 4 == 0:  # through the shape's origin
                p2 = np.array(pos) - (p1 - np.array(pos)) * 0.5
            if i % 9 == 0:  # starting inside
                p1 = np.array(pos) + rng.uniform(-0.1, 0.1, 2)
            fx = b2.Fixture(sh)
            fd, fi = fixture_record(fx)
            hit, normal, fraction = sh.ray_cast(b2.Vec2(*p1), b2.Vec2(*p2), 1.0, xf, 0)
            rows.append((fd, fi, (xf._position.x, xf._position.y, xf._rotation._col1.x, xf._rotation._col1.y),
                         (float(p1[0]), float(p1[1]), float(p2[0]), float(p2[1])), bool(hit),
                         (float(fraction), normal.x if hit else 0.0, normal.y if hit else 0.0)))
    return {"fd": np.array([r[0] for r in rows]), "fi": np.array([r[1] for r in rows]),
            "xf": np.array([r[2] for r in rows]), "seg": np.array([r[3] for r in rows]),
            "hit": np.array([r[4] for r in rows]), "out": np.array([r[5] for r in rows])}


def world_queries(scene, steps, rng, nq=200, nr=200):
    """A reference world after `steps` steps, with World.query_aabb over random boxes (the set of
    fixtures yielded) and World.ray_cast over random segments with every hit continued unclipped
    (ret[-1] = 1): (fixture, point, normal, fraction) of each hit, in the reference's order."""
    world = build_ref_world(scene)
    for _ in range(steps):
        world.step(1.0 / 60, 8, 3)
    state, fidx = snapshot(world)
    bodies = [b for b in world.bodies]
    xs = [b.position.x for b in bodies]
    ys = [b.position.y for b in bodies]
    lo, hi = (min(xs) - 3.0, min(ys) - 3.0), (max(xs) + 3.0, max(ys) + 3.0)
    boxes, qoff, qfix = [], [0], []
    for _ in range(nq):
        c = (float(rng.uniform(lo[0], hi[0])), float(rng.uniform(lo[1], hi[1])))
        e = rng.uniform(0.1, 4.0, 2)
        box = (c[0] - float(e[0]), c[1] - float(e[1]), c[0] + float(e[0]), c[1] + float(e[1]))
        boxes.append(box)
        aabb = b2.AABB(lower_bound=(box[0], box[1]), upper_bound=(box[2], box[3]))
        found = sorted(fidx[id(f)] for f in world.query_aabb(aabb))
        qfix.extend(found)
        qoff.append(len(qfix))
    segs, roff, rfix, rdat = [], [0], [], []
    for _ in range(nr):
        p1 = (float(rng.uniform(lo[0], hi[0])), float(rng.uniform(lo[1], hi[1])))
        p2 = (float(rng.uniform(lo[0], hi[0])), float(rng.uniform(lo[1], hi[1])))
        segs.append(p1 + p2)
        hits = []
        for ret in world.ray_cast(p1, p2):
            f, point, normal, fraction = ret
            hits.append((fidx[id(f)], fraction, point.x, point.y, normal.x, normal.y))
            ret[-1] = 1.0
        hits.sort()
        rfix.extend(h[0] for h in hits)
        rdat.extend(h[1:] for h in hits)
        roff.append(len(rfix))
    out = {"init_" + n: getattr(state, n) for n in L._FIELDS}
    out.update(boxes=np.array(boxes), q_off=np.array(qoff, np.int32), q_fix=np.array(qfix, np.int32),
               segs=np.array(segs), r_off=np.array(roff, np.int32), r_fix=np.array(rfix, np.int32),
               r_dat=np.array(rdat).reshape(-1, 5))
    return out


def main(which=None):
    """Records every fixture, or only those named (pyramid, stacks, mixed, filtered, queries, narrowphase)."""
    want = set(which or ("pyramid", "stacks", "mixed", "filtered", "listeners", "continuous", "joints", "queries",
                         "narrowphase"))
    if "pyramid" in want:
        pyr_ck = {1, 2, 3, 4, 5, 10, 20, 50, 100, 150, 200, 300, 400, 500, 600}
        init, hints, states = record_run(scenes.pyramid(), 600, pyr_ck)
        save_run(os.path.join(HERE, "pyramid_600.npz"), init, hints, states, {"steps": 600})
    if "stacks" in want:
        st_ck = {1, 2, 3, 5, 10, 30, 60, 100, 200, 300, 400, 500, 600}
        init, hints, states = record_run(scenes.stacks(seed=0), 600, st_ck)
        save_run(os.path.join(HERE, "stacks_600.npz"), init, hints, states, {"steps": 600})
    if "mixed" in want:
        mix_ck = {1, 2, 3, 5, 10, 20, 40, 60, 80, 100, 150, 200, 250, 300}
        rng = np.random.default_rng(7)
        init, hints, states = record_run(scenes.mixed_world(rng, 24, half_width=8.0, height=12.0), 300, mix_ck)
        save_run(os.path.join(HERE, "mixed24_300.npz"), init, hints, states, {"steps": 300})
    if "filtered" in want:
        # collision filtering (category / mask / group bits) and a kinematic body
        filt_ck = {1, 2, 3, 5, 10, 20, 40, 60, 80, 100, 120, 150, 200, 250, 300}
        init, hints, states = record_run(scenes.filtered_world(), 300, filt_ck)
        save_run(os.path.join(HERE, "filtered_kinematic_300.npz"), init, hints, states, {"steps": 300})
    if "listeners" in want:
        ck = {1, 2, 5, 10, 20, 40, 60, 61, 80, 100, 110, 111, 112, 130, 150, 151, 170, 200}
        init, hints, states, edit_states, edits, ev_off, ev_i, ev_d = record_listener_run(scenes.filtered_world(), 200,
                                                                                         ck)
        path = os.path.join(HERE, "listeners_filtered_200.npz")
        save_run(path, init, hints, states, {"steps": 200})
        z = dict(np.load(path))
        for s, st in edit_states.items():
            for name in L._FIELDS:
                z["e%d_%s" % (s, name)] = getattr(st, name)
        z.update(edits=edits, ev_off=ev_off, ev_i=ev_i, ev_d=ev_d)
        np.savez_compressed(path, **z)
        print("listener run:", len(ev_i), "calls,", len(edits), "filter edits")
    if "continuous" in want:
        out = record_continuous_run(scenes.pyramid(rows=4), 3)
        np.savez_compressed(os.path.join(HERE, "continuous_pyramid4.npz"), **out)
        print("continuous physics:", [str(out["s%d_error" % s]) for s in (1, 2, 3)], str(out["empty_error"]))
        for name, scene, steps in (("projectiles", scenes.projectiles(), 150), ("pyramid", scenes.pyramid(), 200)):
            stats = []
            ck = {1, 2, 3, 5, 10, 20, 30, 50, 75, 100, 125, 150, 200}
            init, hints, states = record_run(scene, steps, {c for c in ck if c <= steps}, continuous=True, stats=stats)
            path = os.path.join(HERE, "continuous_%s_%d.npz" % (name, steps))
            save_run(path, init, hints, states, {"steps": steps, "continuous": 1})
            z = dict(np.load(path))
            z["toi_stats"] = np.array(stats, np.int64).reshape(-1, 2)
            np.savez_compressed(path, **z)
            print(name, "toi calls", int(z["toi_stats"][:, 0].sum()), "toi events", int(z["toi_stats"][:, 1].sum()),
                  "steps with events", int(np.count_nonzero(z["toi_stats"][:, 1])))
    if "joints" in want:
        ck = {1, 2, 3, 5, 10, 20, 40, 60, 80, 100, 120, 150, 200, 250, 300}
        # the gear scene runs away in the reference (its gear joints' quirks drive the circles to
        # the speed cap; a step near 190 overflows), so it is recorded for 120 steps
        for name, scene, steps in (("bridge", scenes.joint_bridge(), 300), ("machines", scenes.joint_machines(), 300),
                                   ("gears", scenes.joint_gears(), 120)):
            init, hints, states = record_run(scene, steps, {c for c in ck if c <= steps})
            save_run(os.path.join(HERE, "joints_%s_%d.npz" % (name, steps)), init, hints, states, {"steps": steps})
    if "queries" in want:
        np.savez_compressed(os.path.join(HERE, "shape_raycast.npz"), **shape_ray_cases(np.random.default_rng(17)))
        rng = np.random.default_rng(19)
        np.savez_compressed(os.path.join(HERE, "queries_mixed24.npz"),
                            **world_queries(scenes.mixed_world(np.random.default_rng(7), 24, half_width=8.0,
                                                               height=12.0), 120, rng))
        np.savez_compressed(os.path.join(HERE, "queries_filtered.npz"),
                            **world_queries(scenes.filtered_world(), 150, rng))
        print("query and ray-cast fixtures written")
    if "narrowphase" in want:
        cases = narrowphase_cases(np.random.default_rng(11))
        np.savez_compressed(os.path.join(HERE, "narrowphase.npz"), **run_narrowphase(cases))
        print("narrowphase cases:", len(cases))


if __name__ == "__main__":
    main(sys.argv[1:])
