"""Host side of the drop-in API: what is rejected before any device work, and
that the product path has no CPU fallback."""
import os

import numpy as np
import pytest

import pypybox2d_amd as b2
from pypybox2d_amd import _native as N


def _pyramid_world(**kw):
    world = b2.World(gravity=(0, -10), do_sleep=True, **kw)
    ground = world.create_static_body(position=(0, 0))
    ground.create_edge_fixture((-40, 0), (40, 0))
    for i in range(3):
        for j in range(3 - i):
            b = world.create_dynamic_body(position=(-7 + 0.5625 * i + 1.125 * j, 0.75 + i))
            b.create_polygon_fixture(box=(0.5, 0.5), density=5, friction=0.6)
    return world


def test_continuous_physics_with_a_custom_filter_or_sub_stepping_is_rejected_before_device_work():
    world = _pyramid_world(continuous_physics=True)
    world.contact_filter = lambda fa, fb: True
    with pytest.raises(NotImplementedError):
        world.step(1.0 / 60, 8, 3)
    world = _pyramid_world(continuous_physics=True, sub_stepping=True)
    with pytest.raises(NotImplementedError):
        world.step(1.0 / 60, 8, 3)
    assert _pyramid_world()._batch().world_i[0, b2.layout.WI_FLAGS] & b2.layout.WF_CONTINUOUS


def test_sensor_is_rejected():
    world = _pyramid_world(continuous_physics=False)
    body = world.create_dynamic_body(position=(0, 10))
    f = body.create_circle_fixture(radius=0.5, density=1)
    f.sensor = True
    with pytest.raises(NotImplementedError):
        world.step(1.0 / 60, 8, 3)


def test_contact_listener_and_filter_are_set_on_the_contact_manager():
    """contact_manager.py:47-100 and world.py:86-136: the listener's callbacks live on the world's
    contact manager; the contact filter is the world's, World.default_contact_filter unless set."""
    world = _pyramid_world(continuous_physics=False)
    cm = world.contact_manager
    assert cm.contact_listener == [None, None, None, None]
    assert world.contact_filter is b2.World.default_contact_filter and not world._custom_filter()
    f = lambda fa, fb: True  # noqa: E731
    cm.begin_contact = f
    cm.contact_filter = f
    assert cm.contact_listener[0] is f and world.contact_filter is f and world._custom_filter()
    world.contact_filter = None  # the reference's "no filtering"
    assert world._custom_filter() and cm.contact_filter is None
    world2 = b2.World(contact_filter=f)
    assert world2.contact_filter is f


def test_fixture_filter_setters_flag_the_fixtures_contacts():
    """Fixture.refilter (fixture.py:269-285) on the host: the fixture's contacts are flagged and its
    proxy is touched (the world's move buffer)."""
    world = _pyramid_world(continuous_physics=False)
    body = world.bodies[1]
    fx = body.fixtures[0]
    world._con_fix = [(world.bodies[0].fixtures[0], fx), (world.bodies[2].fixtures[0], world.bodies[3].fixtures[0])]
    world._con_i = np.zeros((2, b2.layout.CON_NI), np.int32)
    world._con_d = np.zeros((2, b2.layout.CON_ND))
    world._move = []
    fx.mask_bits = 0x0002
    assert fx.mask_bits == 0x0002
    assert world._con_i[0, b2.layout.CI_FLAGS] & b2.layout.CF_FILTER and not world._con_i[1, b2.layout.CI_FLAGS]
    assert world._move == [fx] and world._refilter_pending
    world._locked = True  # inside a listener callback the change waits for the step's end
    fx.group_index = -2
    assert fx.group_index == -2 and world._deferred_refilter == [fx]


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_no_cpu_fallback_without_a_gpu():
    """The library loads (it was built), but creating a context without a GPU fails loudly."""
    N.load()
    with pytest.raises(N.B2GError):
        N.Engine("fp32")
    world = _pyramid_world(continuous_physics=False)
    with pytest.raises(N.B2GError):
        world.step(1.0 / 60, 8, 3)


def _two_bodies():
    world = b2.World(gravity=(0, -10), continuous_physics=False)
    ground = world.create_static_body(position=(0, 0))
    ground.create_edge_fixture((-10, 0), (10, 0))
    body = world.create_dynamic_body(position=(1, 2), angle=0.3)
    body.create_polygon_fixture(box=(0.5, 0.5), density=2.0)
    return world, ground, body


def test_joint_constructors_check_their_arguments_as_the_reference_does():
    """joints/*.py constructors: missing bodies, bodies of different worlds, a too-short distance,
    a pulley ratio, invalid mouse-joint values, gear joints over other joint types."""
    world, ground, body = _two_bodies()
    with pytest.raises(ValueError):
        world.create_revolute_joint(None, body)
    other = b2.World(continuous_physics=False).create_dynamic_body()
    with pytest.raises(ValueError):
        world.create_weld_joint(body, other, anchor=(0, 0))
    with pytest.raises(ValueError):
        world.create_distance_joint(ground, body, anchor_a=(1, 2), anchor_b=(1, 2.001))
    with pytest.raises(ValueError):
        world.create_pulley_joint(ground, body, ratio=0.0)
    for bad in (dict(target=(float("nan"), 0)), dict(max_force=-1.0), dict(frequency=-1.0),
                dict(damping_ratio=-0.5)):
        with pytest.raises(ValueError):
            world.create_mouse_joint(body, **bad)
    d = world.create_distance_joint(ground, body, anchor_a=(0, 5), anchor_b=(1, 2))
    with pytest.raises(ValueError):
        world.create_gear_joint(d, d)
    assert world.joints == [d] and body.joints == [d] and ground.joints == [d]


def test_joint_properties_follow_the_reference():
    """Anchors through the bodies' transforms, reference angles, local axes, getters and setters
    (setters that wake the bodies do so; a disabled limit clears its impulse)."""
    world, ground, body = _two_bodies()
    r = world.create_revolute_joint(ground, body, anchor=(1, 2.5), lower_angle=-0.5, upper_angle=0.5,
                                    limit_enabled=True, motor_speed=2.0, max_motor_torque=10.0)
    assert abs(r.anchor_a.x - 1.0) < 1e-15 and abs(r.anchor_b.y - 2.5) < 1e-12
    assert abs(r.joint_angle) < 1e-15  # the reference angle is the bodies' angle difference (0.3)
    assert r.limit_enabled and not r.motor_enabled and r.motor_speed == 2.0 and r.max_motor_torque == 10.0
    body.awake = False
    r.motor_speed = 3.0
    assert body.awake and r.motor_speed == 3.0
    r._d[7] = 5.0  # an accumulated limit impulse
    r.limit_enabled = False
    assert not r.limit_enabled and r.get_reaction_torque(60.0) == 0.0
    p = world.create_prismatic_joint(ground, body, anchor=(1, 2), axis=(1, 0))
    assert abs(p.axis.x - 1.0) < 1e-15 and abs(p.joint_translation) < 1e-12
    w = world.create_wheel_joint(ground, body, anchor=(1, 2), axis=(0, 1))
    w.spring_frequency = 9.0  # the reference stores it where its solver never reads it
    assert w.spring_frequency == 9.0
    m = world.create_mouse_joint(body, target=(3, 4), max_force=50.0)
    assert m.body_a is None and m.target.x == 3.0 and m.other_body(body) is None
    m.target = (5, 6)
    assert m.target.y == 6.0
    g = world.create_gear_joint(r, p, ratio=2.0)
    assert g.body_a is body and g.body_b is body and g.ratio == 2.0 and body.joints.count(g) == 2
    with pytest.raises(AttributeError):
        g.get_reaction_force(60.0)
    f = world.create_friction_joint(ground, body, max_force=1.0, max_torque=1.0)
    with pytest.raises(ValueError):
        f.max_force = -1.0


def test_destroying_joints_and_bodies_keeps_the_joint_lists():
    """World.destroy_joint removes the joint from the world and both bodies' lists (twice for a
    gear joint's one body) and wakes the bodies; World.destroy_body destroys the body's joints
    first, calling the joint destruction listener for each."""
    world, ground, body = _two_bodies()
    r = world.create_revolute_joint(ground, body, anchor=(1, 2))
    p = world.create_prismatic_joint(ground, body, anchor=(1, 2), axis=(0, 1))
    g = world.create_gear_joint(r, p, ratio=1.0)
    body.awake = False
    world.destroy_joint(g)
    assert g not in body.joints and body.awake and world.joints == [r, p]
    with pytest.raises(ValueError):
        world.destroy_joint(g)
    seen = []
    world.joint_destruction_listener = seen.append
    world.destroy_body(body)
    assert seen == [r, p] and world.joints == [] and ground.joints == []
    st = world.state()
    assert st.joint_d.shape == (0, 64) and st.body_d.shape[0] == 1
