"""MI355X-native per-step contact pipeline behind a pypybox2d-compatible API.

``World``, ``Body``, ``Fixture`` and the shapes keep pypybox2d's interface;
``World.step`` runs broad phase, narrow phase and the island solver as gfx950
HIP kernels in libb2g.so.  ``BatchedWorld`` steps many independent worlds at
once.  Build the library with ``python -m pypybox2d_amd.build``.
"""
from .common import (AABB, EPSILON, MAX_FLOAT, NUMBER_TYPES, PI, EmptyFixtureError, LockedError, Mat22,
                     PhysicsError, Sweep, Transform, Vec2, clamp, clamp_vector, max_vector, min_vector,
                     scalar_cross)
from .shapes import Circle, Edge, Loop, MassData, Polygon, Shape
from .fixture import Fixture
from .body import Body
from .world import World
from . import joints
from .joints import (DistanceJoint, FrictionJoint, GearJoint, MouseJoint, PrismaticJoint, PulleyJoint, RevoluteJoint,
                     RopeJoint, WeldJoint, WheelJoint)

__all__ = ["World", "Body", "Fixture", "Circle", "Polygon", "Edge", "Loop", "Shape", "MassData", "Vec2", "Mat22",
           "Transform", "AABB", "Sweep", "PhysicsError", "LockedError", "EmptyFixtureError", "PI", "EPSILON",
           "MAX_FLOAT", "NUMBER_TYPES", "scalar_cross", "min_vector", "max_vector", "clamp", "clamp_vector",
           "BatchedWorld", "DistanceJoint", "RevoluteJoint", "FrictionJoint", "PrismaticJoint", "WeldJoint",
           "RopeJoint", "WheelJoint", "MouseJoint", "PulleyJoint", "GearJoint"]


def __getattr__(name):
    if name == "BatchedWorld":
        from .batched import BatchedWorld
        return BatchedWorld
    raise AttributeError(name)
