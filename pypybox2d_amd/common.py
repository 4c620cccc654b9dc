"""2-D math value types of the drop-in API (host side).

Same public behaviour as pypybox2d/common.py (PyVec2 88-225, PyMat22 360-531,
PyTransform 616-704, PyAABB 707-868): value semantics, ``Vec2 * Vec2`` is
the dot product, ``v.cross(s) = (s*y, -s*x)``, ``scalar_cross(s, v) =
(-s*y, s*x)``, and ``Transform.mul_t(Transform)`` keeps the reference's
unrotated translation.  The expressions are written in the reference's
evaluation order, so world-building arithmetic done with these types gives
the same binary64 results.
"""
from __future__ import annotations

import math

from .settings import EPSILON, MAX_FLOAT

PI = math.pi
NUMBER_TYPES = (float, int)


class PhysicsError(Exception):
    pass


class LockedError(PhysicsError):
    pass


class EmptyFixtureError(PhysicsError):
    pass


class Vec2(object):
    """A 2-D column vector with value semantics."""
    __slots__ = ("x", "y")

    def __init__(self, x=0.0, y=0.0):
        self.x, self.y = float(x), float(y)

    def __iter__(self):
        return iter((self.x, self.y))

    def __repr__(self):
        return "Vec2(%g,%g)" % (self.x, self.y)

    def __len__(self):
        return 2

    def __abs__(self):
        return Vec2(abs(self.x), abs(self.y))

    def __neg__(self):
        return Vec2(-self.x, -self.y)

    def __copy__(self):
        return Vec2(self.x, self.y)

    copy = __copy__

    def __add__(self, other):
        ox, oy = other
        return Vec2(self.x + ox, self.y + oy)

    __radd__ = __add__

    def __sub__(self, other):
        return Vec2(self.x - other[0], self.y - other[1])

    def __rsub__(self, other):
        return Vec2(other[0] - self.x, other[1] - self.y)

    def __mul__(self, value):
        if isinstance(value, NUMBER_TYPES):
            return Vec2(value * self.x, value * self.y)
        return self.dot(value)

    __rmul__ = __mul__

    def __truediv__(self, value):
        if isinstance(value, NUMBER_TYPES):
            return Vec2(self.x / value, self.y / value)
        raise ValueError("Ambiguous operation")

    def __eq__(self, other):
        try:
            return self.x == other[0] and self.y == other[1]
        except Exception:
            return False

    def __ne__(self, other):
        return not self.__eq__(other)

    __hash__ = None

    def __getitem__(self, i):
        if i == 0:
            return self.x
        if i == 1:
            return self.y
        raise IndexError("Index must be in (0,1)")

    def __setitem__(self, i, value):
        if i == 0:
            self.x = float(value)
        elif i == 1:
            self.y = float(value)
        else:
            raise IndexError("Index must be in (0,1)")

    @property
    def length(self):
        return math.sqrt(self.x ** 2 + self.y ** 2)

    @property
    def length_squared(self):
        return self.x ** 2 + self.y ** 2

    @property
    def valid(self):
        return not (math.isnan(self.x) or math.isnan(self.y))

    def _scalar_cross(self, value):
        return Vec2(-float(value) * self.y, float(value) * self.x)

    def cross(self, value):
        if isinstance(value, NUMBER_TYPES):
            return Vec2(float(value) * self.y, -float(value) * self.x)
        vx, vy = value
        return self.x * float(vy) - self.y * float(vx)

    def dot(self, value):
        vx, vy = value
        return self.x * float(vx) + self.y * float(vy)

    @property
    def normalized(self):
        length = self.length
        if length < EPSILON:
            return Vec2(*self)
        inv_length = 1.0 / length
        return Vec2(self.x * inv_length, self.y * inv_length)

    def skew(self):
        return Vec2(-self.y, self.x)


def scalar_cross(scalar, vector):
    return Vec2(*vector)._scalar_cross(scalar)


def min_vector(v1, v2):
    return Vec2(min(v1[0], v2[0]), min(v1[1], v2[1]))


def max_vector(v1, v2):
    return Vec2(max(v1[0], v2[0]), max(v1[1], v2[1]))


def clamp(value, low, high):
    return max(low, min(value, high))


def clamp_vector(value, low, high):
    return max_vector(low, min_vector(value, high))


def is_valid_float(x):
    return not math.isnan(x)


class Mat22(object):
    """2x2 matrix, column major."""
    __slots__ = ("col1", "col2")

    def __init__(self, col1=(1, 0), col2=(0, 1), angle=None):
        if angle is not None:
            self.col1, self.col2 = Vec2(), Vec2()
            self.angle = angle
        else:
            self.col1, self.col2 = Vec2(*col1), Vec2(*col2)

    def __copy__(self):
        return Mat22(self.col1, self.col2)

    def __iter__(self):
        return iter((Vec2(*self.col1), Vec2(*self.col2)))

    def __getitem__(self, i):
        return Vec2(*(self.col1, self.col2)[i])

    def __eq__(self, other):
        try:
            return self.col1 == other[0] and self.col2 == other[1]
        except Exception:
            return False

    __hash__ = None

    def __mul__(self, vec):
        v0, v1 = vec
        return Vec2((self.col1.x * v0 + self.col2.x * v1), (self.col1.y * v0 + self.col2.y * v1))

    def mul_t(self, other):
        if isinstance(other, (list, tuple, Vec2)):
            return Vec2(self.col1.dot(other), self.col2.dot(other))
        return Mat22(((self.col1.dot(other.col1)), (self.col2.dot(other.col1))),
                     ((self.col1.dot(other.col2)), (self.col2.dot(other.col2))))

    @property
    def angle(self):
        return math.atan2(self.col1.y, self.col1.x)

    @angle.setter
    def angle(self, radians):
        c = math.cos(radians)
        s = math.sin(radians)
        self.col1.x = c
        self.col2.x = -s
        self.col1.y = s
        self.col2.y = c

    @property
    def inverse(self):
        a, b = self.col1.x, self.col2.x
        c, d = self.col1.y, self.col2.y
        det = a * d - b * c
        if det != 0.0:
            det = 1.0 / det
        return Mat22((det * d, -det * c), (-det * b, det * a))

    def solve(self, vec):
        a, b = self.col1.x, self.col2.x
        c, d = self.col1.y, self.col2.y
        det = a * d - b * c
        if det != 0.0:
            det = 1.0 / det
        return Vec2(det * (d * vec[0] - b * vec[1]), det * (a * vec[1] - c * vec[0]))


class Transform(object):
    """Translation + rotation of a rigid frame."""
    __slots__ = ("_position", "_rotation")

    def __init__(self, position=(0, 0), rotation=((1, 0), (0, 1)), angle=None):
        self._position = Vec2(*position)
        if angle is not None:
            self._rotation = Mat22()
            self._rotation.angle = angle
        else:
            self._rotation = Mat22(*rotation)

    @property
    def position(self):
        return self._position

    @position.setter
    def position(self, position):
        self._position = Vec2(*position)

    @property
    def rotation(self):
        return Mat22(self._rotation.col1, self._rotation.col2)

    @property
    def angle(self):
        return self._rotation.angle

    @angle.setter
    def angle(self, angle):
        self._rotation.angle = angle

    def __copy__(self):
        return Transform(self._position, (self._rotation.col1, self._rotation.col2))

    def __repr__(self):
        return "Transform(position=%s, angle=%g)" % (self._position, self._rotation.angle)

    def __mul__(self, other):
        o0, o1 = other
        col1 = self._rotation.col1
        col2 = self._rotation.col2
        return Vec2(self._position.x + col1.x * o0 + col2.x * o1,
                    self._position.y + col1.y * o0 + col2.y * o1)

    def mul_t(self, other):
        if isinstance(other, (list, tuple, Vec2)):
            return self._rotation.mul_t(Vec2(*other) - self._position)
        if isinstance(other, Transform):
            # same translation as the reference (not rotated), common.py:670-673
            t = Transform(other.position - self._position)
            t._rotation = self._rotation.mul_t(other._rotation)
            return t
        raise TypeError


class AABB(object):
    """An axis-aligned bounding box."""
    __slots__ = ("_lower_bound", "_upper_bound")

    def __init__(self, lower_bound=(0, 0), upper_bound=(0, 0)):
        self._lower_bound = Vec2(*lower_bound)
        self._upper_bound = Vec2(*upper_bound)

    def __repr__(self):
        return "AABB(lower_bound=%s, upper_bound=%s)" % (self._lower_bound, self._upper_bound)

    @property
    def lower_bound(self):
        return Vec2(*self._lower_bound)

    @lower_bound.setter
    def lower_bound(self, value):
        self._lower_bound = Vec2(*value)

    @property
    def upper_bound(self):
        return Vec2(*self._upper_bound)

    @upper_bound.setter
    def upper_bound(self, value):
        self._upper_bound = Vec2(*value)

    def __iter__(self):
        return iter((Vec2(*self._lower_bound), Vec2(*self._upper_bound)))

    def __eq__(self, other):
        try:
            return self._lower_bound == other[0] and self._upper_bound == other[1]
        except Exception:
            return False

    __hash__ = None

    def __getitem__(self, i):
        return Vec2(*(self._lower_bound, self._upper_bound)[i])

    def __copy__(self):
        return AABB(self._lower_bound, self._upper_bound)

    def __add__(self, other):
        return AABB(min_vector(self._lower_bound, other._lower_bound),
                    max_vector(self._upper_bound, other._upper_bound))

    @property
    def center(self):
        return 0.5 * (self._lower_bound + self._upper_bound)

    @property
    def extents(self):
        return 0.5 * (self._upper_bound - self._lower_bound)

    @property
    def perimeter(self):
        wx = self._upper_bound.x - self._lower_bound.x
        wy = self._upper_bound.y - self._lower_bound.y
        return 2.0 * (wx + wy)

    def test_overlap(self, b):
        d1 = b._lower_bound - self._upper_bound
        d2 = self._lower_bound - b._upper_bound
        if d1.x > 0.0 or d1.y > 0.0:
            return False
        if d2.x > 0.0 or d2.y > 0.0:
            return False
        return True

    def contains(self, aabb):
        return ((self._lower_bound.x <= aabb._lower_bound.x) and (self._lower_bound.y <= aabb._lower_bound.y)
                and (aabb._upper_bound.x <= self._upper_bound.x) and (aabb._upper_bound.y <= self._upper_bound.y))


class Sweep(object):
    """Centre-of-mass motion of a body over a step."""
    __slots__ = ("local_center", "c0", "c", "a0", "a", "alpha0")

    def __init__(self, local_center=(0, 0), c0=(0, 0), c=(0, 0), a0=0.0, a=0.0, alpha0=0.0):
        self.local_center = Vec2(*local_center)
        self.c0, self.c = Vec2(*c0), Vec2(*c)
        self.a0, self.a = a0, a
        self.alpha0 = alpha0

    def __copy__(self):
        return Sweep(self.local_center, self.c0, self.c, self.a0, self.a, self.alpha0)


__all__ = ["PhysicsError", "LockedError", "EmptyFixtureError", "PI", "EPSILON", "MAX_FLOAT", "NUMBER_TYPES",
           "Vec2", "Mat22", "Transform", "AABB", "Sweep", "scalar_cross", "min_vector", "max_vector", "clamp",
           "clamp_vector", "is_valid_float"]
