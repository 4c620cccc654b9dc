"""The oracle's joints (pypybox2d/joints/, island.py:107-148) against recorded reference runs
and against what the constraints fix.

Fixtures (tests/golden/make_golden.py, the reference's pure-Python flavour):
  * joints_bridge_300.npz: the testbed bridge, 30 planks on revolute joints, shapes dropped
    onto it;
  * joints_machines_300.npz: every joint type but gear in one world: pendulums between revolute
    limits (both limits reached), a motor, equal limits, a prismatic elevator (limits and
    motor) and slide, a welded cantilever, a chain under a rope joint, rigid and soft distance
    joints (one with collide_connected), a pulley, a car on wheel joints, a friction joint and a
    mouse joint;
  * joints_gears_120.npz: the testbed gears (revolute-revolute and revolute-prismatic gear
    joints; the scene runs away in the reference after about 180 steps).
The oracle replays each run with the recorded contact-creation order and must match every
checkpoint (bodies, contacts, joint records) bit for bit.  The closed-form checks hold the
constraint equations themselves.
"""
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN
from oracle import oracle as O
from pypybox2d_amd import builder, layout as L, scenes

RUNS = ["joints_bridge_300.npz", "joints_machines_300.npz", "joints_gears_120.npz"]


@pytest.mark.parametrize("run", RUNS)
def test_oracle_replays_joint_reference_runs_bit_exact(run):
    z = np.load(os.path.join(GOLDEN, run))
    init = L.load_batch(z, "init_")
    assert init.n_joints > 5
    o = O.Oracle(init)
    hoff, hp = z["hint_off"], z["hint_pairs"]
    cks = set(int(s) for s in z["checkpoints"])
    for s in range(1, int(z["meta_steps"]) + 1):
        o.set_hint(0, hp[hoff[s - 1]:hoff[s]])
        assert o.step() == 0, s
        if s in cks:
            st = o.state()
            for name in ("body_d", "fix_d", "con_d", "con_i", "move", "joint_d", "joint_i"):
                got, want = getattr(st, name), z["s%d_%s" % (s, name)]
                assert got.shape == want.shape and np.array_equal(got, want), (s, name)


def test_machines_run_reaches_every_limit_state():
    """The recorded run drives the revolute pendulums to both limits, holds one at equal limits,
    and moves the prismatic elevator through all three of its states."""
    z = np.load(os.path.join(GOLDEN, "joints_machines_300.npz"))
    o = O.Oracle(L.load_batch(z, "init_"))
    hoff, hp = z["hint_off"], z["hint_pairs"]
    seen = {}
    for s in range(1, 301):
        o.set_hint(0, hp[hoff[s - 1]:hoff[s]])
        o.step()
        st = o.state()
        for k in range(st.joint_i.shape[0]):
            seen.setdefault(k, set()).add(int(st.joint_i[k, L.JI_STATE]))
    ji = o.state().joint_i
    rev = [k for k in range(len(ji)) if ji[k, L.JI_TYPE] == L.JT_REVOLUTE and ji[k, L.JI_FLAGS] & L.JF_LIMIT]
    assert any({L.LIMIT_AT_LOWER, L.LIMIT_AT_UPPER} <= seen[k] for k in rev)
    assert any(seen[k] == {L.LIMIT_EQUAL} for k in rev)
    pri = [k for k in range(len(ji)) if ji[k, L.JI_TYPE] == L.JT_PRISMATIC and ji[k, L.JI_FLAGS] & L.JF_LIMIT]
    assert any({L.LIMIT_INACTIVE, L.LIMIT_AT_LOWER, L.LIMIT_AT_UPPER} <= seen[k] for k in pri)


def _world(bodies, joints, gravity=(0.0, -10.0)):
    return {"gravity": gravity, "do_sleep": False, "warm_starting": True, "bodies": bodies, "joints": joints}


def _dyn(pos, shape, density=1.0, **kw):
    return scenes._body("dynamic", pos, [scenes._fix(shape, density=density)], **kw)


GROUND = scenes._body("static", (0.0, 0.0), [scenes._fix(("edge", (-50.0, -20.0), (50.0, -20.0)))])


def _world_point(st, b, local):
    c, s = st.body_d[b, L.BD_XF_C], st.body_d[b, L.BD_XF_S]
    return np.array([st.body_d[b, L.BD_XF_PX] + c * local[0] - s * local[1],
                     st.body_d[b, L.BD_XF_PY] + s * local[0] + c * local[1]])


def test_revolute_anchor_and_distance_length_hold():
    """A double pendulum (revolute joints) and a rigid distance joint: after 200 steps of
    swinging the joint anchors coincide and the distance stays its length, within the position
    solver's slop (0.005 m), while the bodies have moved by metres."""
    scene = _world([GROUND, _dyn((1.0, 0.0), ("box", 1.0, 0.1)), _dyn((3.0, 0.0), ("box", 1.0, 0.1)),
                    _dyn((6.0, 0.0), ("circle", 0.3))],
                   [scenes._joint("revolute", 0, 1, anchor=(0.0, 0.0)),
                    scenes._joint("revolute", 1, 2, anchor=(2.0, 0.0)),
                    scenes._joint("distance", 2, 3, anchor_a=(4.0, 0.0), anchor_b=(6.0, 0.0))])
    b = builder.scene_batch(scene)
    o = O.Oracle(b)
    for _ in range(200):
        assert o.step() == 0
    st = o.state()
    jd = st.joint_d
    for k, (ba, bb) in enumerate(((0, 1), (1, 2))):
        pa = _world_point(st, ba, jd[k, L.JD_LA_X:L.JD_LA_X + 2])
        pb = _world_point(st, bb, jd[k, L.JD_LB_X:L.JD_LB_X + 2])
        assert np.linalg.norm(pa - pb) < 0.005, k
    pa = _world_point(st, 2, jd[2, L.JD_LA_X:L.JD_LA_X + 2])
    pb = _world_point(st, 3, jd[2, L.JD_LB_X:L.JD_LB_X + 2])
    assert abs(np.linalg.norm(pa - pb) - 2.0) < 0.005
    assert st.body_d[3, L.BD_C_Y] < -2.0  # it swung down


def test_prismatic_keeps_its_axis_and_weld_its_angle():
    """A box on a prismatic joint along the unit axis (1, 1)/sqrt(2) falls along that line (no
    rotation, no drift off the axis); a box welded to a swinging plank keeps the reference
    angle.  (The reference does not normalise the axis it is given, prismatic.py:86-89, and its
    effective masses assume a unit axis.)"""
    h = math.sqrt(0.5)
    scene = _world([GROUND, _dyn((0.0, 0.0), ("box", 0.3, 0.3)), _dyn((5.0, 0.0), ("box", 0.5, 0.2)),
                    _dyn((6.0, 0.0), ("box", 0.5, 0.2), angle=0.4)],
                   [scenes._joint("prismatic", 0, 1, anchor=(0.0, 0.0), axis=(h, h)),
                    scenes._joint("revolute", 0, 2, anchor=(4.5, 0.0)),
                    scenes._joint("weld", 2, 3, anchor=(5.5, 0.0))])
    o = O.Oracle(builder.scene_batch(scene))
    for _ in range(120):
        assert o.step() == 0
    st = o.state()
    x, y, a = st.body_d[1, L.BD_C_X], st.body_d[1, L.BD_C_Y], st.body_d[1, L.BD_A]
    assert x < -3.0 and abs(x - y) < 0.005 and abs(a) < 2.0 / 180.0 * math.pi
    rel = st.body_d[3, L.BD_A] - st.body_d[2, L.BD_A]
    assert abs(rel - 0.4) < 2.0 / 180.0 * math.pi and abs(st.body_d[2, L.BD_A]) > 0.5


def test_revolute_motor_reaches_its_speed():
    """A wheel on a revolute motor (speed 3 rad/s, ample torque, no gravity) spins at exactly
    the motor speed after the first step; with a tiny torque limit it only accelerates by
    max_torque * dt / I per step."""
    def run(torque):
        scene = _world([GROUND, _dyn((0.0, 5.0), ("circle", 1.0))],
                       [scenes._joint("revolute", 0, 1, anchor=(0.0, 5.0), motor_enabled=True, motor_speed=3.0,
                                      max_motor_torque=torque)], gravity=(0.0, 0.0))
        b = builder.scene_batch(scene)
        o = O.Oracle(b)
        o.step()
        return o.state().body_d[1, L.BD_W], b.body_d[1, L.BD_I]
    w, _ = run(1000.0)
    assert abs(w - 3.0) < 1e-12
    w, inertia = run(0.01)
    assert abs(w - 0.01 / 60.0 / inertia) < 1e-12


def test_joined_bodies_do_not_collide_unless_allowed():
    """Body._should_collide (body.py:920-926): the bridge's neighbouring planks overlap at their
    ends, yet no contact joins two planks pinned together; two overlapping circles on a distance
    joint get a contact only with collide_connected."""
    o = O.Oracle(builder.scene_batch(scenes.joint_bridge()))
    for _ in range(5):
        o.step()
    st = o.state()
    planks = set(range(1, 31))
    pairs = {(int(st.fix_i[a, L.FI_BODY]), int(st.fix_i[b, L.FI_BODY])) for a, b in st.con_i[:, :2]}
    assert not any(a in planks and b in planks and abs(a - b) == 1 for a, b in pairs)
    for cc in (False, True):
        scene = _world([GROUND, _dyn((0.0, 0.0), ("circle", 0.5)), _dyn((0.6, 0.0), ("circle", 0.5))],
                       [scenes._joint("distance", 1, 2, anchor_a=(0.0, 0.0), anchor_b=(0.6, 0.0),
                                      collide_connected=cc)], gravity=(0.0, 0.0))
        o = O.Oracle(builder.scene_batch(scene))
        o.step()
        assert o.state().con_i.shape[0] == (1 if cc else 0), cc


@pytest.mark.parametrize("run,scene", [("joints_bridge_300.npz", scenes.joint_bridge),
                                       ("joints_machines_300.npz", scenes.joint_machines),
                                       ("joints_gears_120.npz", scenes.joint_gears)])
def test_drop_in_joint_constructors_build_the_reference_records(run, scene):
    """World.create_*_joint on the drop-in API (anchors through get_local_point, reference
    angles, local axes, pulley constants, gear coordinates) gives the joint records the
    reference's objects hold after construction, bit for bit."""
    z = np.load(os.path.join(GOLDEN, run))
    ref = L.load_batch(z, "init_")
    got = builder.scene_batch(scene())
    for name in L.ALL_FIELDS:
        assert np.array_equal(getattr(got, name), getattr(ref, name)), name
