"""Contacts as the host sees them, and the world's contact manager.

``Contact`` is the object the contact listener receives and ``World.contacts``
lists (pypybox2d/contact.py:49-160): one per fixture pair for as long as the
pair's contact exists, holding the manifold of the device's record at the
listener call (or after the last step).  ``Manifold`` / ``ManifoldPoint`` /
``WorldManifold`` follow contact_util.py:49-245.  ``ContactManager`` is the
world's ``contact_manager`` (contact_manager.py:32-100): the listener's four
callbacks and the contact filter are set on it.

The manifold and flags are the device's; ``Contact.enabled`` may be set to
False inside begin_contact or pre_solve, which removes the contact from this
step's solve (the world applies it between collide and the solve).  Other
changes made inside a callback are refused (the world is locked, and Body
setters raise LockedError) since the step's remaining work already runs on the
device.
"""
from __future__ import annotations

from copy import copy

from . import layout as L
from .common import Vec2
from .settings import EPSILON

__all__ = ("ManifoldPoint", "Manifold", "WorldManifold", "Contact", "ContactManager")

EPSILON_SQR = EPSILON * EPSILON


class ManifoldPoint(object):
    """One contact point (contact_util.py:49-103): local point, accumulated impulses, id."""
    __slots__ = ("local_point", "normal_impulse", "tangent_impulse", "id")
    VERTEX, FACE = 0, 1

    def __init__(self, local_point=(0, 0), normal_impulse=0.0, tangent_impulse=0.0, id=None):
        self.local_point = Vec2(*local_point)
        self.normal_impulse = normal_impulse
        self.tangent_impulse = tangent_impulse
        self.id = (0, 0, 0, 0) if id is None else tuple(id)

    def __copy__(self):
        return ManifoldPoint(self.local_point, self.normal_impulse, self.tangent_impulse, self.id)

    def __repr__(self):
        return "ManifoldPoint(local_point=%s, normal_impulse=%g, tangent_impulse=%g, id=%s)" % (
            self.local_point, self.normal_impulse, self.tangent_impulse, self.id)

    index_a = property(lambda self: self.id[0])
    index_b = property(lambda self: self.id[1])
    type_a = property(lambda self: self.id[2])
    type_b = property(lambda self: self.id[3])


class Manifold(object):
    """A contact manifold (contact_util.py:139-193)."""
    __slots__ = ("points", "local_normal", "local_point", "type", "point_count")
    CIRCLES, FACE_A, FACE_B = 0, 1, 2
    NULL_STATE, ADD_STATE, PERSIST_STATE, REMOVE_STATE = 0, 1, 2, 3

    def __init__(self, points=None, local_normal=(0, 0), local_point=(0, 0), type=CIRCLES, point_count=0):
        self.points = [ManifoldPoint(), ManifoldPoint()] if points is None else [copy(p) for p in points]
        self.local_normal = Vec2(*local_normal)
        self.local_point = Vec2(*local_point)
        self.type = type
        self.point_count = point_count

    def __copy__(self):
        return Manifold(self.points, self.local_normal, self.local_point, self.type, self.point_count)

    @classmethod
    def from_record(cls, d, i) -> "Manifold":
        """The 12 float64 and 4 int32 manifold fields of an event record (b2g_layout.h)."""
        pts = [ManifoldPoint((d[4], d[5]), float(d[6]), float(d[7]), L.unpack_id(i[2])),
               ManifoldPoint((d[8], d[9]), float(d[10]), float(d[11]), L.unpack_id(i[3]))]
        return cls(pts, (d[0], d[1]), (d[2], d[3]), int(i[0]), int(i[1]))

    @classmethod
    def from_contact_row(cls, cd, ci) -> "Manifold":
        """The manifold fields of a contact record of the batch layout."""
        pts = [ManifoldPoint((cd[L.CD_P0_X], cd[L.CD_P0_Y]), float(cd[L.CD_P0_NI]), float(cd[L.CD_P0_TI]),
                             L.unpack_id(ci[L.CI_ID0])),
               ManifoldPoint((cd[L.CD_P1_X], cd[L.CD_P1_Y]), float(cd[L.CD_P1_NI]), float(cd[L.CD_P1_TI]),
                             L.unpack_id(ci[L.CI_ID1]))]
        return cls(pts, (cd[L.CD_LN_X], cd[L.CD_LN_Y]), (cd[L.CD_LP_X], cd[L.CD_LP_Y]), int(ci[L.CI_MTYPE]),
                   int(ci[L.CI_PC]))

    @property
    def used_points(self):
        return self.points[:self.point_count]

    def get_point_states(self, manifold2):
        """The points' states from this manifold to manifold2 (contact_util.py:169-193): for this
        one's points PERSIST or REMOVE, for manifold2's ADD or PERSIST, by point id."""
        ids1 = [p.id for p in self.used_points]
        ids2 = [p.id for p in manifold2.used_points]
        state1 = [self.PERSIST_STATE if i in ids2 else self.REMOVE_STATE for i in ids1]
        state2 = [self.PERSIST_STATE if i in ids1 else self.ADD_STATE for i in ids2]
        return state1, state2


class WorldManifold(object):
    """The manifold's normal and points in world coordinates for the given transforms and shape radii
    (contact_util.py:195-245): the points lie midway between the two surfaces."""
    __slots__ = ("normal", "points")

    def __init__(self, manifold, xf_a, radius_a, xf_b, radius_b):
        self.normal = Vec2()
        self.points = []
        if manifold.point_count == 0:
            return
        if manifold.type == Manifold.CIRCLES:
            pa = xf_a * manifold.local_point
            pb = xf_b * manifold.points[0].local_point
            n = Vec2(1.0, 0.0)
            d = pb - pa
            if d.x ** 2 + d.y ** 2 > EPSILON_SQR:
                n = d.normalized
            self.points = [0.5 * ((pa + radius_a * n) + (pb - radius_b * n))]
            self.normal = n
        elif manifold.type == Manifold.FACE_A:
            n = xf_a._rotation * manifold.local_normal
            plane = xf_a * manifold.local_point
            for mp in manifold.used_points:
                clip = xf_b * mp.local_point
                ca = clip + (radius_a - (clip - plane).dot(n)) * n
                self.points.append(0.5 * (ca + (clip - radius_b * n)))
            self.normal = n
        elif manifold.type == Manifold.FACE_B:
            n = xf_b._rotation * manifold.local_normal
            plane = xf_b * manifold.local_point
            for mp in manifold.used_points:
                clip = xf_a * mp.local_point
                cb = clip + (radius_b - (clip - plane).dot(n)) * n
                self.points.append(0.5 * ((clip - radius_a * n) + cb))
            self.normal = -n  # from A to B


class Contact(object):
    """A contact between two fixtures (contact.py:49-160), as the device last reported it."""

    def __init__(self, world, fixture_a, fixture_b):
        self._world = world
        self._fixture_a, self._fixture_b = fixture_a, fixture_b
        self._touching_flag = False
        self._enabled_flag = True
        self._filter_flag = False
        self._friction = (fixture_a._friction * fixture_b._friction) ** 0.5
        self._restitution = max(fixture_a._restitution, fixture_b._restitution)
        self._manifold = Manifold()

    def _load_event(self, ei, ed):
        f = int(ei[L.EI_FLAGS])
        self._touching_flag = bool(f & L.CF_TOUCHING)
        self._enabled_flag = bool(f & L.CF_ENABLED)
        self._manifold = Manifold.from_record(ed[:12], ei[L.EI_MTYPE:L.EI_MTYPE + 4])

    def _load_row(self, cd, ci):
        f = int(ci[L.CI_FLAGS])
        self._touching_flag = bool(f & L.CF_TOUCHING)
        self._enabled_flag = bool(f & L.CF_ENABLED)
        self._filter_flag = bool(f & L.CF_FILTER)
        self._friction = float(cd[L.CD_FRICTION])
        self._restitution = float(cd[L.CD_RESTITUTION])
        self._manifold = Manifold.from_contact_row(cd, ci)

    def __repr__(self):
        return "Contact(touching=%s, point_count=%d)" % (self._touching_flag, self._manifold.point_count)

    @property
    def manifold(self):
        return self._manifold

    @property
    def world_manifold(self):
        return WorldManifold(self._manifold, self.body_a._xf, self.shape_a.radius, self.body_b._xf,
                             self.shape_b.radius)

    @property
    def touching(self):
        return self._touching_flag

    @property
    def enabled(self):
        return self._enabled_flag

    @enabled.setter
    def enabled(self, enabled):
        """contact.py:116-127: inside begin_contact or pre_solve, False keeps the contact out of this
        step's solve; the next step's update enables it again."""
        self._enabled_flag = bool(enabled)

    fixture_a = property(lambda self: self._fixture_a)
    fixture_b = property(lambda self: self._fixture_b)
    body_a = property(lambda self: self._fixture_a._body)
    body_b = property(lambda self: self._fixture_b._body)
    shape_a = property(lambda self: self._fixture_a._shape)
    shape_b = property(lambda self: self._fixture_b._shape)
    friction = property(lambda self: self._friction)
    restitution = property(lambda self: self._restitution)

    def other_body(self, body):
        if body is self._fixture_a._body:
            return self._fixture_b._body
        return self._fixture_a._body


class ContactManager(object):
    """The world's contact manager (contact_manager.py:32-100): the contact listener's callbacks
    begin_contact(contact), end_contact(contact), pre_solve(contact, old_manifold) and
    post_solve(contact, impulses), and the contact filter, which is the world's."""

    def __init__(self, world):
        self._world = world
        self.begin_contact = None
        self.end_contact = None
        self.pre_solve = None
        self.post_solve = None

    @property
    def contact_filter(self):
        return self._world.contact_filter

    @contact_filter.setter
    def contact_filter(self, fcn):
        self._world.contact_filter = fcn

    @property
    def contact_listener(self):
        return [self.begin_contact, self.end_contact, self.pre_solve, self.post_solve]

    @property
    def contacts(self):
        return self._world.contacts
