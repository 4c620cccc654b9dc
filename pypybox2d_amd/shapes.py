"""Collision shapes of the drop-in API (host side).

Circle, Polygon, Edge and Loop with the construction-time geometry of
pypybox2d/shapes.py: polygon normals ``normalize(edge.cross(1))`` and the
convexity check (shapes.py:442-482), boxes (484-501), mass data (191-201,
336-419, 720-723) and AABBs (178-189, 316-334, 702-718).  Shapes are copied
when a fixture is made (fixture.py:67); a polygon copy rebuilds its normals
from the vertices (shapes.py:250-251), as the reference does.
Loop shapes can be built but a world holding one cannot be stepped on the
device path (b2g_upload reports B2G_E_UNSUPPORTED).
"""
from __future__ import annotations

import math
import operator
from copy import copy
from functools import reduce

from .common import AABB, PI, Transform, Vec2, max_vector, min_vector
from .settings import EPSILON, EPSILON_SQR, MAX_POLYGON_VERTICES, POLYGON_RADIUS

__all__ = ("MassData", "Shape", "Circle", "Polygon", "Edge", "Loop")


class MassData(object):
    def __init__(self, mass=0.0, center=(0, 0), I=0.0):
        self.mass = mass
        self.center = Vec2(*center)
        self.I = I

    def __repr__(self):
        return "MassData(mass=%g, center=%s, I=%g)" % (self.mass, self.center, self.I)


class Shape(object):
    radius = 0.0
    _type = -1

    @property
    def child_count(self):
        return 1


class Circle(Shape):
    _type = 0

    def __init__(self, radius=0.2, position=(0, 0)):
        self.position = Vec2(*position)
        self.radius = radius

    def __copy__(self):
        return Circle(self.radius, self.position)

    def __repr__(self):
        return "Circle(radius=%g, position=%s)" % (self.radius, self.position)

    def compute_aabb(self, xf, child_index=0):
        p = xf._position + xf._rotation * self.position
        rad = (self.radius, self.radius)
        return AABB(p - rad, p + rad)

    def compute_mass(self, density):
        md = MassData()
        md.mass = density * PI * self.radius ** 2
        md.center = copy(self.position)
        md.I = md.mass * (0.5 * self.radius ** 2 + self.position * self.position)
        return md

    def test_point(self, xf, p):
        center = xf._position + xf._rotation * self.position
        d = Vec2(*p) - center
        return (d * d) <= self.radius ** 2


class Polygon(Shape):
    _type = 1
    CHECK_VERTICES = True

    def __init__(self, vertices=None, box=None):
        self.radius = POLYGON_RADIUS
        self._vertices = []
        self._normals = []
        self._centroid = Vec2()
        self.centroid = Vec2()
        if box:
            self.set_as_box(*box)
        elif vertices is not None:
            self.vertices = [Vec2(*v) for v in vertices]

    def __copy__(self):
        return Polygon(self._vertices)

    def __repr__(self):
        return "Polygon(%s)" % self._vertices

    @property
    def vertices(self):
        return self._vertices

    @vertices.setter
    def vertices(self, vertices):
        vertices = [Vec2(*v) for v in vertices]
        count = len(vertices)
        if count < 3:
            raise ValueError("Polygons are >= 3 vertices")
        normals = []
        for i in range(count):
            edge = vertices[(i + 1) % count] - vertices[i]
            if edge.length_squared <= EPSILON_SQR:
                raise ValueError("Edges are too short")
            normals.append(edge.cross(1.0).normalized)
        if Polygon.CHECK_VERTICES:
            for i1 in range(count):
                i2 = (i1 + 1) % count
                edge = vertices[i2] - vertices[i1]
                for j in range(count):
                    if j in (i1, i2):
                        continue
                    if edge.cross(vertices[j] - vertices[i1]) <= 0.0:
                        raise ValueError("polygon must be convex with a counter-clockwise winding")
        self._normals = normals
        self._vertices = [Vec2(*v) for v in vertices]
        self.centroid = self.compute_centroid()

    @property
    def normals(self):
        return list(self._normals)

    def set_as_box(self, hx, hy, center=(0, 0), angle=0.0):
        self._vertices = [Vec2(-hx, -hy), Vec2(hx, -hy), Vec2(hx, hy), Vec2(-hx, hy)]
        self._normals = [Vec2(0, -1), Vec2(1, 0), Vec2(0, 1), Vec2(-1, 0)]
        self.centroid = Vec2(*center)
        if center is not None:
            xf = Transform(center, angle=angle)
            self._vertices = [xf * v for v in self._vertices]
            self._normals = [xf * v for v in self._normals]

    def compute_centroid(self, vertices=None):
        vertices = self._vertices if vertices is None else vertices
        if len(vertices) < 3:
            raise ValueError("Need >= 3 vertices to compute the centroid")
        p_ref = Vec2()
        c = Vec2()
        area = 0.0
        inv3 = 1.0 / 3.0
        count = len(vertices)
        for i, v in enumerate(vertices):
            p1, p2, p3 = p_ref, v, vertices[(i + 1) % count]
            D = (p2 - p1).cross(p3 - p1)
            triangle_area = 0.5 * D
            area += triangle_area
            c += triangle_area * inv3 * (p1 + p2 + p3)
        if area <= EPSILON:
            raise ValueError("Area of the polygon is too small to get a centroid")
        return c * (1.0 / area)

    def compute_aabb(self, xf, child_index=0):
        lower = xf * self._vertices[0]
        upper = lower
        for vertex in self._vertices:
            v = xf * vertex
            lower = min_vector(lower, v)
            upper = max_vector(upper, v)
        r = (self.radius, self.radius)
        return AABB(lower - r, upper + r)

    def compute_mass(self, density):
        verts = self._vertices
        if len(verts) < 3:
            raise ValueError("Polygons are >= 3 vertices")
        md = MassData()
        center = Vec2()
        area = 0.0
        I = 0.0
        s = reduce(operator.add, verts) / len(verts)
        inv3 = 1.0 / 3.0
        count = len(verts)
        for i, vertex in enumerate(verts):
            e1 = vertex - s
            e2 = verts[(i + 1) % count] - s
            D = e1.cross(e2)
            triangle_area = 0.5 * D
            area += triangle_area
            center += triangle_area * inv3 * (e1 + e2)
            ex1, ey1 = e1
            ex2, ey2 = e2
            intx2 = ex1 * ex1 + ex2 * ex1 + ex2 * ex2
            inty2 = ey1 * ey1 + ey2 * ey1 + ey2 * ey2
            I += (0.25 * inv3 * D) * (intx2 + inty2)
        md.mass = density * area
        if not area > EPSILON:
            raise ValueError("polygon area is too small")
        center /= area
        md.center = center + s
        md.I = density * I
        md.I += md.mass * (md.center.length_squared - center.length_squared)
        return md

    def test_point(self, xf, p):
        p_local = xf._rotation.mul_t(Vec2(*p) - xf._position)
        for normal, vertex in zip(self._normals, self._vertices):
            if normal.dot(p_local - vertex) > 0.0:
                return False
        return True


class Edge(Shape):
    _type = 3

    def __init__(self, v1=None, v2=None, v0=None, v3=None):
        self.radius = POLYGON_RADIUS
        self._vertex1 = Vec2(*v1)
        self._vertex2 = Vec2(*v2)
        self._vertex0 = Vec2(*v0) if v0 is not None else None
        self._vertex3 = Vec2(*v3) if v3 is not None else None

    def __copy__(self):
        return Edge(self._vertex1, self._vertex2, self._vertex0, self._vertex3)

    def __repr__(self):
        return "Edge(vertex1=%s, vertex2=%s, [vertex0=%s, vertex3=%s])" % (
            self._vertex1, self._vertex2, self._vertex0, self._vertex3)

    vertex0 = property(lambda self: None if self._vertex0 is None else Vec2(*self._vertex0))
    vertex1 = property(lambda self: Vec2(*self._vertex1))
    vertex2 = property(lambda self: Vec2(*self._vertex2))
    vertex3 = property(lambda self: None if self._vertex3 is None else Vec2(*self._vertex3))

    def compute_aabb(self, xf, child_index=0):
        v1 = xf * self._vertex1
        v2 = xf * self._vertex2
        r = (self.radius, self.radius)
        return AABB(min_vector(v1, v2) - r, max_vector(v1, v2) + r)

    def compute_mass(self, density):
        return MassData(mass=0.0, center=0.5 * (self._vertex1 + self._vertex2), I=0.0)

    def test_point(self, xf, p):
        return False


class Loop(Shape):
    """Closed chain of edges (kept for API compatibility; not steppable on the device path)."""
    _type = 2

    def __init__(self, vertices, *args):
        if len(vertices) == 2 and isinstance(vertices[0], (int, float)):
            vertices = [vertices] + list(args)
        if len(vertices) < 2:
            raise ValueError("Must be >= 2 vertices")
        self.radius = POLYGON_RADIUS
        self._vertices = [Vec2(*v) for v in vertices]

    def __copy__(self):
        return Loop(self._vertices)

    @property
    def child_count(self):
        return len(self._vertices)

    def compute_mass(self, density):
        return MassData()

    def compute_aabb(self, xf, child_index=0):
        i1 = child_index
        i2 = (child_index + 1) % len(self._vertices)
        v1 = xf * self._vertices[i1]
        v2 = xf * self._vertices[i2]
        return AABB(min_vector(v1, v2), max_vector(v1, v2))


def check_device_shape(shape):
    """Raise if the device path cannot hold this shape."""
    if isinstance(shape, Loop):
        raise NotImplementedError("Loop shapes are outside the device contact pipeline")
    if isinstance(shape, Polygon) and len(shape._vertices) > MAX_POLYGON_VERTICES:
        raise NotImplementedError("polygons are limited to %d vertices on the device" % MAX_POLYGON_VERTICES)
