"""Host side of the drop-in API: what is rejected before any device work, and
that the product path has no CPU fallback."""
import os

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


def test_continuous_physics_is_rejected_before_device_work():
    world = _pyramid_world(continuous_physics=True)
    with pytest.raises(NotImplementedError):
        world.step(1.0 / 60, 8, 3)


def test_sensor_is_rejected():
    world = _pyramid_world(continuous_physics=False)
    body = world.create_dynamic_body(position=(0, 10))
    f = body.create_circle_fixture(radius=0.5, density=1)
    f.sensor = True
    with pytest.raises(NotImplementedError):
        world.step(1.0 / 60, 8, 3)


def test_contact_listener_is_rejected():
    world = _pyramid_world(continuous_physics=False)
    world.begin_contact = lambda contact: None
    with pytest.raises(NotImplementedError):
        world.step(1.0 / 60, 8, 3)


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_no_cpu_fallback_without_a_gpu():
    """The library loads (it was built), but creating a context without a GPU fails loudly."""
    N.load()
    with pytest.raises(N.B2GError):
        N.Engine("fp32")
    world = _pyramid_world(continuous_physics=False)
    with pytest.raises(N.B2GError):
        world.step(1.0 / 60, 8, 3)
