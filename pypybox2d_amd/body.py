"""Rigid body of the drop-in API (host side).

Constructor, mass computation and the public accessors follow
pypybox2d/body.py:39-116 (Body.__init__), 185-214 (attach_fixture), 288-308
(_attached_to_world) and 606-660 (reset_mass_data), including the velocity
shift ``v += w x (c_new - c_old)`` each time the centre of mass moves.

While a world is being stepped on the GPU the authoritative state lives on the
device; reading a body pulls the state back lazily (one download per step),
and writing one marks the world for re-upload before the next step.
"""
from __future__ import annotations

from copy import copy

from .common import LockedError, Sweep, Transform, Vec2, is_valid_float, scalar_cross
from .fixture import Fixture
from .shapes import MassData, Shape

__all__ = ("Body",)


class Body(object):
    STATIC = 0
    DYNAMIC = 1
    KINEMATIC = 2

    def __init__(self, world=None, position=(0, 0), angle=0.0, linear_velocity=(0, 0), angular_velocity=0.0,
                 linear_damping=0.0, angular_damping=0.0, allow_sleep=True, awake=True, fixed_rotation=False,
                 bullet=False, type=STATIC, active=True, gravity_scale=1.0, check_args=True, shapes=(),
                 fixtures=(), shape_fixture=None, user_data=None, copy_fixtures=True):
        position = Vec2(*position)
        linear_velocity = Vec2(*linear_velocity)
        if check_args:
            if not position.valid:
                raise ValueError("Invalid position")
            if not linear_velocity.valid:
                raise ValueError("Invalid linear velocity")
            if not is_valid_float(angle):
                raise ValueError("Invalid angle")
            if not is_valid_float(angular_velocity):
                raise ValueError("Invalid angular velocity")
            if type not in (Body.STATIC, Body.DYNAMIC, Body.KINEMATIC):
                raise ValueError("Invalid body type")
        self._xf = Transform(position=position, angle=angle)
        self._sweep = Sweep()
        self._force = Vec2()
        self._torque = 0.0
        self._I = 0.0
        self._invI = 0.0
        self._sleep_time = 0.0
        self.fixtures = []
        self.joints = []
        self.user_data = user_data
        self._world = world
        self._index = -1
        self._linear_velocity = linear_velocity
        self._angular_velocity = float(angular_velocity)
        self._linear_damping = float(linear_damping)
        self._angular_damping = float(angular_damping)
        self._gravity_scale = float(gravity_scale)
        self._awake = awake
        self._allow_sleep = bool(allow_sleep)
        self._fixed_rotation = bool(fixed_rotation)
        self._bullet = bullet
        self._type = type
        self._active = active
        self._sweep.local_center = Vec2()
        self._sweep.a0 = self._sweep.a = angle
        self._sweep.c0 = self._xf * self._sweep.local_center
        self._sweep.c = copy(self._sweep.c0)
        if self._type == Body.DYNAMIC:
            self._mass = 1.0
            self._inv_mass = 1.0
        else:
            self._mass = 0.0
            self._inv_mass = 0.0
        if shapes is not None:
            self.create_fixtures(shapes, shape_fixture)
        if fixtures is not None:
            if not isinstance(fixtures, (list, tuple)):
                fixtures = [fixtures]
            for fixture in fixtures:
                self.attach_fixture(copy(fixture) if copy_fixtures else fixture)

    def __repr__(self):
        return "Body(type=%d, position=%s, angle=%g)" % (self._type, self.position, self.angle)

    # ---- device synchronisation ----
    def _pull(self):
        if self._world is not None:
            self._world._pull()

    def _modified(self):
        if self._world is not None:
            if self._world._locked:
                raise LockedError
            self._world._pull()
            self._world._host_dirty = True

    # ---- fixtures ----
    def create_fixture(self, shape, friction=0.2, restitution=0.0, density=0.0, category_bits=0x0001,
                       mask_bits=0xFFFF, group_index=0, sensor=False, user_data=None):
        if self._world and self._world._locked:
            raise LockedError
        fixture = Fixture(shape, friction, restitution, density, category_bits, mask_bits, group_index, sensor,
                          user_data, body=self)
        self.attach_fixture(fixture)
        return fixture

    def create_fixtures(self, shape_list, fixture=None, friction=0.2, restitution=0.0, density=0.0,
                        category_bits=0x0001, mask_bits=0xFFFF, group_index=0, sensor=False):
        if isinstance(shape_list, Shape):
            shape_list = [shape_list]
        if fixture is not None:
            friction, restitution, density = fixture.friction, fixture.restitution, fixture.density
            category_bits, mask_bits = fixture.category_bits, fixture.mask_bits
            group_index, sensor = fixture.group_index, fixture.sensor
        for shape in shape_list:
            self.create_fixture(shape, friction, restitution, density, category_bits, mask_bits, group_index,
                                sensor)

    def create_circle_fixture(self, radius=0.2, position=(0, 0), **kwargs):
        from .shapes import Circle
        return self.create_fixture(Circle(radius, position), **kwargs)

    def create_polygon_fixture(self, vertices=None, box=None, **kwargs):
        from .shapes import Polygon
        return self.create_fixture(Polygon(vertices, box), **kwargs)

    def create_edge_fixture(self, v1=None, v2=None, v0=None, v3=None, **kwargs):
        from .shapes import Edge
        return self.create_fixture(Edge(v1, v2, v0, v3), **kwargs)

    def attach_fixture(self, fixture):
        if fixture in self.fixtures:
            raise ValueError("Fixture is already attached to this body")
        if self._world is not None:
            self._world._pull()
        if fixture._body is not None and fixture._body is not self:
            fixture._body.destroy_fixture(fixture)
        fixture._body = self
        if fixture._shape is None:
            raise ValueError("Cannot attach empty fixture to a body")
        self.fixtures.append(fixture)
        if fixture._density > 0.0:
            self.reset_mass_data()
        if self._world is not None:
            self._world._fixture_added(self, fixture)

    def destroy_fixture(self, fixture):
        if self._world and self._world._locked:
            raise LockedError
        if fixture._body is not self or fixture not in self.fixtures:
            raise ValueError("This fixture does not belong to this body")
        if self._world is not None:
            self._world._pull()
            self._world._fixture_removed(self, fixture)
        self.fixtures.remove(fixture)
        fixture._body = None
        self.reset_mass_data()

    # ---- mass ----
    def reset_mass_data(self):
        self._mass = 0.0
        self._inv_mass = 0.0
        self._I = 0.0
        self._invI = 0.0
        self._sweep.local_center = Vec2()
        if self._type in (Body.STATIC, Body.KINEMATIC):
            self._sweep.c = self._xf.position
            self._sweep.c0 = copy(self._sweep.c)
            return
        center = Vec2()
        for fixture in self.fixtures:
            if fixture._density == 0.0:
                continue
            md = fixture.mass_data
            self._mass += md.mass
            center += md.mass * md.center
            self._I += md.I
        if self._mass > 0.0:
            self._inv_mass = 1.0 / self._mass
            center *= self._inv_mass
        else:
            self._mass = 1.0
            self._inv_mass = 1.0
        if self._I > 0.0 and not self._fixed_rotation:
            self._I -= self._mass * center.length_squared
            if not self._I > 0.0:
                raise ValueError("non-positive rotational inertia about the centre of mass")
            self._invI = 1.0 / self._I
        else:
            self._I = 0.0
            self._invI = 0.0
        old_center = self._sweep.c
        self._sweep.local_center = center
        self._sweep.c0 = self._xf * self._sweep.local_center
        self._sweep.c = copy(self._sweep.c0)
        self._linear_velocity += scalar_cross(self._angular_velocity, self._sweep.c - old_center)
        if self._world is not None:
            self._world._host_dirty = True

    @property
    def mass(self):
        return self._mass

    @property
    def inertia(self):
        return self._I + self._mass * self._sweep.local_center.length_squared

    @property
    def mass_data(self):
        return MassData(mass=self._mass, I=self._I + self._mass * (self._sweep.local_center * self._sweep.local_center),
                        center=self._sweep.local_center)

    @mass_data.setter
    def mass_data(self, md):
        self._modified()
        if self._type != Body.DYNAMIC:
            return
        self._inv_mass = 0.0
        self._I = 0.0
        self._invI = 0.0
        self._mass = md.mass
        if self._mass <= 0.0:
            self._mass = 1.0
        self._inv_mass = 1.0 / self._mass
        if md.I > 0.0 and not self._fixed_rotation:
            self._I = md.I - self._mass * (md.center * md.center)
            self._invI = 1.0 / self._I
        old_center = self._sweep.c
        self._sweep.local_center = copy(md.center)
        self._sweep.c0 = self._xf * self._sweep.local_center
        self._sweep.c = copy(self._sweep.c0)
        self._linear_velocity += scalar_cross(self.angular_velocity, self._sweep.c - old_center)

    # ---- state ----
    @property
    def type(self):
        return self._type

    @property
    def dynamic(self):
        return self._type == Body.DYNAMIC

    @property
    def static(self):
        return self._type == Body.STATIC

    @property
    def kinematic(self):
        return self._type == Body.KINEMATIC

    @property
    def transform(self):
        self._pull()
        return copy(self._xf)

    @property
    def position(self):
        self._pull()
        return Vec2(*self._xf.position)

    @position.setter
    def position(self, pos):
        self._modified()
        self._xf.position = pos
        self._transform_updated()

    @property
    def angle(self):
        self._pull()
        return self._xf.angle

    @angle.setter
    def angle(self, angle):
        self._modified()
        self._xf.angle = angle
        self._transform_updated()

    def _transform_updated(self):
        self._sweep.c0 = self._xf * self._sweep.local_center
        self._sweep.c = copy(self._sweep.c0)
        self._sweep.a0 = self._sweep.a = self._xf.angle
        if self._world is not None:
            self._world._transform_changed(self)

    @property
    def world(self):
        return self._world

    @property
    def world_center(self):
        self._pull()
        return copy(self._sweep.c)

    @property
    def local_center(self):
        return copy(self._sweep.local_center)

    @property
    def linear_velocity(self):
        self._pull()
        return copy(self._linear_velocity)

    @linear_velocity.setter
    def linear_velocity(self, v):
        if self._type == Body.STATIC:
            return
        self._modified()
        v = Vec2(*v)
        if v.dot(v) > 0.0:
            self.awake = True
        self._linear_velocity = v

    @property
    def angular_velocity(self):
        self._pull()
        return self._angular_velocity

    @angular_velocity.setter
    def angular_velocity(self, w):
        if self._type == Body.STATIC:
            return
        self._modified()
        if w * w > 0.0:
            self.awake = True
        self._angular_velocity = float(w)

    def apply_force(self, force, point):
        if self._type != Body.DYNAMIC:
            return
        self._modified()
        if not self._awake:
            self.awake = True
        self._force += force
        self._torque += (Vec2(*point) - self._sweep.c).cross(force)

    def apply_force_to_center(self, force):
        if self._type != Body.DYNAMIC:
            return
        self._modified()
        if not self._awake:
            self.awake = True
        self._force += force

    def apply_torque(self, torque):
        if self._type != Body.DYNAMIC:
            return
        self._modified()
        if not self._awake:
            self.awake = True
        self._torque += torque

    def apply_linear_impulse(self, impulse, point):
        if self._type != Body.DYNAMIC:
            return
        self._modified()
        if not self._awake:
            self.awake = True
        self._linear_velocity += self._inv_mass * Vec2(*impulse)
        self._angular_velocity += self._invI * (Vec2(*point) - self._sweep.c).cross(impulse)

    def apply_angular_impulse(self, impulse):
        if self._type != Body.DYNAMIC:
            return
        self._modified()
        if not self._awake:
            self.awake = True
        self._angular_velocity += self._invI * impulse

    @property
    def awake(self):
        self._pull()
        return self._awake

    @awake.setter
    def awake(self, awake):
        self._modified()
        if awake:
            if not self._awake:
                self._sleep_time = 0.0
        else:
            self._sleep_time = 0.0
            self._linear_velocity = Vec2()
            self._angular_velocity = 0.0
            self._force = Vec2()
            self._torque = 0.0
        self._awake = awake

    @property
    def sleeping(self):
        return not self.awake

    @property
    def allow_sleep(self):
        return self._allow_sleep

    @allow_sleep.setter
    def allow_sleep(self, allow_sleep):
        self._modified()
        self._allow_sleep = allow_sleep
        if not allow_sleep:
            self.awake = True

    @property
    def active(self):
        return self._active

    @property
    def bullet(self):
        return self._bullet

    @property
    def fixed_rotation(self):
        return self._fixed_rotation

    @property
    def linear_damping(self):
        return self._linear_damping

    @linear_damping.setter
    def linear_damping(self, d):
        self._modified()
        self._linear_damping = float(d)

    @property
    def angular_damping(self):
        return self._angular_damping

    @angular_damping.setter
    def angular_damping(self, d):
        self._modified()
        self._angular_damping = float(d)

    @property
    def gravity_scale(self):
        return self._gravity_scale

    @gravity_scale.setter
    def gravity_scale(self, s):
        self._modified()
        self._gravity_scale = float(s)

    @property
    def contacts(self):
        """Contacts touching this body (read-only views, in creation order)."""
        if self._world is None:
            return []
        return [c for c in self._world.contacts if c.body_a is self or c.body_b is self]

    def _should_collide(self, other):
        """body.py:910-928: at least one body dynamic, and no joint between the two that keeps
        them from colliding."""
        if self._type != Body.DYNAMIC and other._type != Body.DYNAMIC:
            return False
        for joint in self.joints:
            if (joint.body_a is self and joint.body_b is other) or (joint.body_b is self and joint.body_a is other):
                if not joint.collide_connected:
                    return False
        return True

    def get_world_point(self, local_point):
        self._pull()
        return self._xf * local_point

    def get_world_vector(self, local_vector):
        self._pull()
        return self._xf._rotation * local_vector

    def get_local_point(self, world_point):
        self._pull()
        return self._xf.mul_t(world_point)

    def get_local_vector(self, world_vector):
        self._pull()
        return self._xf._rotation.mul_t(world_vector)

    def get_linear_velocity_from_world_point(self, world_point):
        self._pull()
        return self._linear_velocity + scalar_cross(self._angular_velocity, Vec2(*world_point) - self._sweep.c)
