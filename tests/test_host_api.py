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
