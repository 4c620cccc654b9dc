"""World: the drop-in entry point (pypybox2d/world.py:64-354).

``World.step(dt, velocity_iterations, position_iterations)`` runs the contact
pipeline on the GPU through libb2g.so (one world, one context).  World
building stays on the host and reproduces the reference's construction
arithmetic: bodies get their mass data and initial fat AABBs exactly as
World.add_body / Body._attached_to_world / Fixture._create_proxies compute
them (world.py:183-224, body.py:288-308, fixture.py:250-261,
dynamictree.py:115-134).

``World.query_aabb`` and ``World.ray_cast`` run on an LBVH the device builds
over the fat AABBs of the current state.

Outside the device path, and therefore refused with NotImplementedError when a
world holding them is stepped: continuous physics (TOI), joints, controllers,
sensors, loop shapes, contact listeners and custom contact filters.
"""
from __future__ import annotations

from copy import copy
from typing import List, Optional

import numpy as np

from . import layout as L
from .body import Body
from .common import LockedError, Vec2
from .settings import AABB_EXTENSION
from .shapes import Circle, Edge, Loop, Polygon, check_device_shape

__all__ = ("World", "ContactView", "TimeStep")


class TimeStep(object):
    __slots__ = ("dt", "inv_dt", "dt_ratio", "vel_iters", "pos_iters", "warm_starting")

    def __init__(self, dt=0.0, inv_dt=0.0, dt_ratio=0.0, vel_iters=0, pos_iters=0, warm_starting=True):
        self.dt, self.inv_dt, self.dt_ratio = float(dt), float(inv_dt), float(dt_ratio)
        self.vel_iters, self.pos_iters, self.warm_starting = int(vel_iters), int(pos_iters), bool(warm_starting)


class ManifoldPointView(object):
    __slots__ = ("local_point", "normal_impulse", "tangent_impulse", "id")

    def __init__(self, lp, ni, ti, id_):
        self.local_point, self.normal_impulse, self.tangent_impulse, self.id = lp, ni, ti, id_


class ManifoldView(object):
    __slots__ = ("type", "point_count", "local_normal", "local_point", "points")

    @property
    def used_points(self):
        return self.points[:self.point_count]


class ContactView(object):
    """Read-only view of one contact (contact.py:49-160) after the last step."""

    def __init__(self, world, k, fa, fb):
        self._fixture_a, self._fixture_b = fa, fb
        cd, ci = world._con_d[k], world._con_i[k]
        self._touching = bool(ci[L.CI_FLAGS] & L.CF_TOUCHING)
        self._enabled = bool(ci[L.CI_FLAGS] & L.CF_ENABLED)
        self.friction = float(cd[L.CD_FRICTION])
        self.restitution = float(cd[L.CD_RESTITUTION])
        m = ManifoldView()
        m.type = int(ci[L.CI_MTYPE])
        m.point_count = int(ci[L.CI_PC])
        m.local_normal = Vec2(cd[L.CD_LN_X], cd[L.CD_LN_Y])
        m.local_point = Vec2(cd[L.CD_LP_X], cd[L.CD_LP_Y])
        m.points = [ManifoldPointView(Vec2(cd[L.CD_P0_X], cd[L.CD_P0_Y]), float(cd[L.CD_P0_NI]),
                                      float(cd[L.CD_P0_TI]), L.unpack_id(ci[L.CI_ID0])),
                    ManifoldPointView(Vec2(cd[L.CD_P1_X], cd[L.CD_P1_Y]), float(cd[L.CD_P1_NI]),
                                      float(cd[L.CD_P1_TI]), L.unpack_id(ci[L.CI_ID1]))]
        self.manifold = m

    fixture_a = property(lambda self: self._fixture_a)
    fixture_b = property(lambda self: self._fixture_b)
    body_a = property(lambda self: self._fixture_a._body)
    body_b = property(lambda self: self._fixture_b._body)
    touching = property(lambda self: self._touching)
    enabled = property(lambda self: self._enabled)

    def other_body(self, body):
        return self.body_b if body is self.body_a else self.body_a


def body_record(b: Body):
    """One body as a (float64 row, int32 row) of the batch layout."""
    d = np.zeros(L.BODY_ND)
    i = np.zeros(L.BODY_NI, np.int32)
    xf, sw = b._xf, b._sweep
    d[L.BD_XF_PX], d[L.BD_XF_PY] = xf._position.x, xf._position.y
    d[L.BD_XF_C], d[L.BD_XF_S] = xf._rotation.col1.x, xf._rotation.col1.y
    d[L.BD_LC_X], d[L.BD_LC_Y] = sw.local_center.x, sw.local_center.y
    d[L.BD_C0_X], d[L.BD_C0_Y] = sw.c0.x, sw.c0.y
    d[L.BD_C_X], d[L.BD_C_Y] = sw.c.x, sw.c.y
    d[L.BD_A0], d[L.BD_A] = sw.a0, sw.a
    d[L.BD_V_X], d[L.BD_V_Y] = b._linear_velocity.x, b._linear_velocity.y
    d[L.BD_W] = b._angular_velocity
    d[L.BD_F_X], d[L.BD_F_Y] = b._force.x, b._force.y
    d[L.BD_TORQUE] = b._torque
    d[L.BD_MASS], d[L.BD_INV_MASS], d[L.BD_I], d[L.BD_INV_I] = b._mass, b._inv_mass, b._I, b._invI
    d[L.BD_GRAV_SCALE] = b._gravity_scale
    d[L.BD_LIN_DAMP], d[L.BD_ANG_DAMP] = b._linear_damping, b._angular_damping
    d[L.BD_SLEEP_TIME] = b._sleep_time
    i[L.BI_TYPE] = b._type
    i[L.BI_FLAGS] = ((L.BF_AWAKE if b._awake else 0) | (L.BF_ALLOW_SLEEP if b._allow_sleep else 0)
                     | (L.BF_FIXED_ROT if b._fixed_rotation else 0) | (L.BF_ACTIVE if b._active else 0)
                     | (L.BF_BULLET if b._bullet else 0))
    i[L.BI_ISLAND] = -1
    return d, i


def fixture_record(f, body_index: int):
    d = np.zeros(L.FIX_ND)
    i = np.zeros(L.FIX_NI, np.int32)
    sh = f._shape
    d[L.FD_FRICTION], d[L.FD_RESTITUTION], d[L.FD_DENSITY] = f._friction, f._restitution, f._density
    d[L.FD_RADIUS] = sh.radius
    i[L.FI_BODY] = body_index
    i[L.FI_SHAPE] = sh._type
    i[L.FI_CATEGORY], i[L.FI_MASK], i[L.FI_GROUP] = f._category_bits, f._mask_bits, f._group_index
    i[L.FI_SENSOR] = int(bool(f._sensor))
    if isinstance(sh, Circle):
        d[L.FD_CIRCLE_PX], d[L.FD_CIRCLE_PY] = sh.position.x, sh.position.y
    elif isinstance(sh, Polygon):
        i[L.FI_NVERTS] = len(sh._vertices)
        for k, (v, n) in enumerate(zip(sh._vertices, sh._normals)):
            d[L.FD_VERTS + 2 * k], d[L.FD_VERTS + 2 * k + 1] = v.x, v.y
            d[L.FD_NORMALS + 2 * k], d[L.FD_NORMALS + 2 * k + 1] = n.x, n.y
        d[L.FD_CENTROID_X], d[L.FD_CENTROID_Y] = sh._centroid.x, sh._centroid.y
    elif isinstance(sh, Edge):
        d[L.FD_EDGE_V1X], d[L.FD_EDGE_V1Y] = sh._vertex1.x, sh._vertex1.y
        d[L.FD_EDGE_V2X], d[L.FD_EDGE_V2Y] = sh._vertex2.x, sh._vertex2.y
        flags = 0
        if sh._vertex0 is not None:
            d[L.FD_EDGE_V0X], d[L.FD_EDGE_V0Y] = sh._vertex0.x, sh._vertex0.y
            flags |= L.EF_HAS_V0
        if sh._vertex3 is not None:
            d[L.FD_EDGE_V3X], d[L.FD_EDGE_V3Y] = sh._vertex3.x, sh._vertex3.y
            flags |= L.EF_HAS_V3
        i[L.FI_EDGE_FLAGS] = flags
    elif isinstance(sh, Loop):
        i[L.FI_SHAPE] = L.SHAPE_LOOP
    lo, hi = f._fat
    d[L.FD_FAT_LX], d[L.FD_FAT_LY], d[L.FD_FAT_UX], d[L.FD_FAT_UY] = lo.x, lo.y, hi.x, hi.y
    return d, i


def fat_aabb(shape, xf):
    """DynamicTree.create_proxy's fat AABB: tight AABB grown by AABB_EXTENSION."""
    aabb = shape.compute_aabb(xf, 0)
    r = Vec2(AABB_EXTENSION, AABB_EXTENSION)
    return (aabb.lower_bound - r, aabb.upper_bound + r)


class World(object):
    def __init__(self, gravity=(0, -10), do_sleep=True, warm_starting=True, continuous_physics=True,
                 sub_stepping=False, auto_clear_forces=True, joint_destruction_listener=None,
                 fixture_destruction_listener=None, contact_filter=None, precision="fp64", device=0,
                 contact_factor=8.0):
        self._gravity = Vec2(*gravity)
        self.bodies: List[Body] = []
        self.joints = []
        self.controllers = []
        self._warm_starting = warm_starting
        self._continuous_physics = continuous_physics
        self._sub_stepping = sub_stepping
        self._allow_sleep = do_sleep
        self._auto_clear_forces = auto_clear_forces
        self.joint_destruction_listener = joint_destruction_listener
        self.fixture_destruction_listener = fixture_destruction_listener
        self._contact_filter = contact_filter
        self.begin_contact = self.end_contact = self.pre_solve = self.post_solve = None
        self._inv_dt0 = 0.0
        self._new_fixture = False
        self._locked = False
        self._precision = precision
        self._device = device
        self._contact_factor = contact_factor
        self._stepper = None        # the device batch (a Stepper: grows its contact slots when needed)
        self._device_ahead = False  # device holds state newer than the Body objects
        self._host_dirty = True     # Body objects changed since the last upload
        self._move: List = []       # broad-phase move buffer (fixtures)
        self._con_d = np.zeros((0, L.CON_ND))
        self._con_i = np.zeros((0, L.CON_NI), np.int32)
        self._con_fix: List = []    # (fixture_a, fixture_b) per contact row

    # ---- properties ----
    @property
    def gravity(self):
        return Vec2(*self._gravity)

    @gravity.setter
    def gravity(self, g):
        self._pull()
        self._gravity = Vec2(*g)
        self._host_dirty = True

    @property
    def locked(self):
        return self._locked

    @property
    def warm_starting(self):
        return self._warm_starting

    @warm_starting.setter
    def warm_starting(self, v):
        self._pull()
        self._warm_starting = bool(v)
        self._host_dirty = True

    @property
    def allow_sleep(self):
        return self._allow_sleep

    @allow_sleep.setter
    def allow_sleep(self, v):
        self._pull()
        self._allow_sleep = bool(v)
        self._host_dirty = True

    @property
    def continuous_physics(self):
        return self._continuous_physics

    @continuous_physics.setter
    def continuous_physics(self, v):
        self._continuous_physics = bool(v)

    @property
    def auto_clear_forces(self):
        return self._auto_clear_forces

    @auto_clear_forces.setter
    def auto_clear_forces(self, v):
        self._pull()
        self._auto_clear_forces = bool(v)
        self._host_dirty = True

    @property
    def contact_filter(self):
        return self._contact_filter

    @contact_filter.setter
    def contact_filter(self, fcn):
        self._contact_filter = fcn

    # ---- bodies ----
    def add_body(self, body: Body) -> Body:
        if self._locked:
            raise LockedError("Cannot create a body while simulating")
        self._pull()
        self.bodies.append(body)
        body._world = self
        if body.fixtures:
            reset = False
            for f in body.fixtures:
                if body._active:
                    f._fat = fat_aabb(f._shape, body._xf)
                    self._move.append(f)
                if f._density > 0.0:
                    reset = True
            self._new_fixture = True
            if reset:
                body.reset_mass_data()
        self._host_dirty = True
        return body

    def create_body(self, body=None, **kwargs):
        if self._locked:
            raise LockedError("Cannot create a body while simulating")
        body = copy(body) if body is not None else Body(**kwargs)
        return self.add_body(body)

    def create_dynamic_body(self, body=None, **kwargs):
        return self.create_body(body, type=Body.DYNAMIC, **kwargs)

    def create_static_body(self, body=None, **kwargs):
        return self.create_body(body, type=Body.STATIC, **kwargs)

    def create_kinematic_body(self, body=None, **kwargs):
        return self.create_body(body, type=Body.KINEMATIC, **kwargs)

    def destroy_body(self, body: Body):
        if self._locked:
            raise LockedError("Cannot remove a body while simulating")
        if body not in self.bodies:
            raise ValueError("Body not in this world")
        self._pull()
        fixtures = set(id(f) for f in body.fixtures)
        self._drop_contacts(lambda fa, fb: id(fa) in fixtures or id(fb) in fixtures)
        self._move = [f for f in self._move if id(f) not in fixtures]
        self.bodies.remove(body)
        body._world = None
        self._host_dirty = True

    # hooks used by Body / Fixture
    def _fixture_added(self, body, fixture):
        if body._active:
            fixture._fat = fat_aabb(fixture._shape, body._xf)
            self._move.append(fixture)
        self._new_fixture = True
        self._host_dirty = True

    def _fixture_removed(self, body, fixture):
        self._drop_contacts(lambda fa, fb: fa is fixture or fb is fixture)
        self._move = [f for f in self._move if f is not fixture]
        self._host_dirty = True

    def _topology_changed(self):
        self._pull()
        self._host_dirty = True

    def _transform_changed(self, body):
        """Body.transform setter: re-sync the proxies (body.py:386-400, fixture.py:287-298).
        The new pairs are found at the start of the next step."""
        for f in body.fixtures:
            aabb = f._shape.compute_aabb(body._xf, 0)
            lo, hi = f._fat
            if lo.x <= aabb._lower_bound.x and lo.y <= aabb._lower_bound.y and \
                    aabb._upper_bound.x <= hi.x and aabb._upper_bound.y <= hi.y:
                continue
            r = Vec2(AABB_EXTENSION, AABB_EXTENSION)
            f._fat = (aabb.lower_bound - r, aabb.upper_bound + r)
            if f not in self._move:
                self._move.append(f)
        self._new_fixture = True
        self._host_dirty = True

    def _drop_contacts(self, pred):
        keep = [k for k, (fa, fb) in enumerate(self._con_fix) if not pred(fa, fb)]
        self._con_d = self._con_d[keep]
        self._con_i = self._con_i[keep]
        self._con_fix = [self._con_fix[k] for k in keep]

    # ---- device state ----
    def _fixture_list(self):
        return [f for b in self.bodies for f in b.fixtures]

    def _batch(self) -> L.BatchArrays:
        fixtures = self._fixture_list()
        fidx = {id(f): k for k, f in enumerate(fixtures)}
        nb, nf = len(self.bodies), len(fixtures)
        bd = np.zeros((nb, L.BODY_ND))
        bi = np.zeros((nb, L.BODY_NI), np.int32)
        bidx = {}
        for k, b in enumerate(self.bodies):
            bd[k], bi[k] = body_record(b)
            bidx[id(b)] = k
        fd = np.zeros((nf, L.FIX_ND))
        fi = np.zeros((nf, L.FIX_NI), np.int32)
        for k, f in enumerate(fixtures):
            fd[k], fi[k] = fixture_record(f, bidx[id(f._body)])
        ci = self._con_i.copy()
        for k, (fa, fb) in enumerate(self._con_fix):
            ci[k, L.CI_FA] = fidx[id(fa)]
            ci[k, L.CI_FB] = fidx[id(fb)]
        move = np.array([fidx[id(f)] for f in self._move], np.int32)
        wd = np.zeros((1, L.WORLD_ND))
        wd[0, L.WD_GX], wd[0, L.WD_GY] = self._gravity.x, self._gravity.y
        wd[0, L.WD_INV_DT0] = self._inv_dt0
        wi = np.zeros((1, L.WORLD_NI), np.int32)
        wi[0, L.WI_FLAGS] = ((L.WF_ALLOW_SLEEP if self._allow_sleep else 0)
                             | (L.WF_WARM_STARTING if self._warm_starting else 0)
                             | (L.WF_NEW_FIXTURE if self._new_fixture else 0)
                             | (0 if self._auto_clear_forces else L.WF_KEEP_FORCES))
        return L.BatchArrays(np.array([0, nb], np.int32), np.array([0, nf], np.int32),
                             np.array([0, len(ci)], np.int32), np.array([0, len(move)], np.int32),
                             wd, wi, bd, bi, fd, fi, self._con_d.copy(), ci, move)

    def _check_supported(self):
        if self._continuous_physics:
            raise NotImplementedError("continuous physics (TOI) is not part of the device contact pipeline; "
                                      "create the world with continuous_physics=False")
        if self.joints or self.controllers:
            raise NotImplementedError("joints and controllers are outside the device contact pipeline")
        if self._contact_filter is not None or any((self.begin_contact, self.end_contact, self.pre_solve,
                                                    self.post_solve)):
            raise NotImplementedError("contact listeners and custom contact filters are not supported")
        for b in self.bodies:
            for f in b.fixtures:
                if f._sensor:
                    raise NotImplementedError("sensor fixtures are not supported")
                check_device_shape(f._shape)

    def _push(self):
        from .stepper import Stepper
        self._check_supported()
        if self._stepper is None:
            self._stepper = Stepper(self._precision, self._device, self._contact_factor)
        self._stepper.upload(self._batch())
        self._host_dirty = False

    def _pull(self):
        """Copy the device state into the Body objects (after a step)."""
        if not self._device_ahead:
            return
        st = self._stepper.download()
        self._device_ahead = False
        fixtures = self._fixture_list()
        for k, b in enumerate(self.bodies):
            d, i = st.body_d[k], st.body_i[k]
            b._xf = _transform(d)
            b._sweep.local_center = Vec2(d[L.BD_LC_X], d[L.BD_LC_Y])
            b._sweep.c0 = Vec2(d[L.BD_C0_X], d[L.BD_C0_Y])
            b._sweep.c = Vec2(d[L.BD_C_X], d[L.BD_C_Y])
            b._sweep.a0, b._sweep.a = float(d[L.BD_A0]), float(d[L.BD_A])
            b._linear_velocity = Vec2(d[L.BD_V_X], d[L.BD_V_Y])
            b._angular_velocity = float(d[L.BD_W])
            b._force = Vec2(d[L.BD_F_X], d[L.BD_F_Y])
            b._torque = float(d[L.BD_TORQUE])
            b._sleep_time = float(d[L.BD_SLEEP_TIME])
            b._awake = bool(i[L.BI_FLAGS] & L.BF_AWAKE)
        for k, f in enumerate(fixtures):
            r = st.fix_d[k]
            f._fat = (Vec2(r[L.FD_FAT_LX], r[L.FD_FAT_LY]), Vec2(r[L.FD_FAT_UX], r[L.FD_FAT_UY]))
        self._con_d = st.con_d.copy()
        self._con_i = st.con_i.copy()
        self._con_fix = [(fixtures[a], fixtures[b]) for a, b in zip(st.con_i[:, L.CI_FA], st.con_i[:, L.CI_FB])]
        self._move = [fixtures[k] for k in st.move]
        self._inv_dt0 = float(st.world_d[0, L.WD_INV_DT0])
        self._new_fixture = bool(st.world_i[0, L.WI_FLAGS] & L.WF_NEW_FIXTURE)

    def step(self, dt, vel_iters, pos_iters):
        """One World.step (world.py:307-354) on the GPU."""
        if self._locked:
            raise LockedError
        if self._host_dirty or self._stepper is None:
            self._pull()
            self._push()
        self._locked = True
        try:
            # steps, then checks the device's error flags: a capacity error grows the contact
            # slots and replays from the last synchronised state; non-finite state raises
            self._stepper.step(float(dt), int(vel_iters), int(pos_iters))
            self._device_ahead = True
            self._new_fixture = False
            if dt > 0.0:
                self._inv_dt0 = 1.0 / dt
        finally:
            self._locked = False

    # ---- queries (world.py:371-440), on the device's LBVH over the fat AABBs ----
    def _device_state(self):
        """The device batch of this world, current with the Body objects."""
        if self._host_dirty or self._stepper is None:
            self._pull()
            self._push()
        return self._stepper.engine

    def query_aabb(self, aabb, callback=None):
        """Fixtures whose fat AABB (the broad-phase proxy) overlaps `aabb` (World.query_aabb,
        world.py:371-384).  As in the reference, the result is a generator: without a callback it
        yields the fixtures; with one, iterating it calls the callback with each fixture until the
        callback returns a false value, and yields nothing.  Fixtures come in fixture order (the
        reference's order is that of its tree, which is not part of the result)."""
        lo, hi = (aabb.lower_bound, aabb.upper_bound) if hasattr(aabb, "lower_bound") else aabb
        box = np.array([[lo[0], lo[1], hi[0], hi[1]]], np.float64)
        rows = self._device_state().query_aabb(np.zeros(1, np.int32), box)
        fixtures = self._fixture_list()
        found = [fixtures[int(f)] for f in rows[:, 1]]
        return self._query_hits(found, callback)

    @staticmethod
    def _query_hits(found, callback):
        for f in found:
            if callback is None:
                yield f
            elif not callback(f):
                break

    def ray_cast(self, point1, point2, callback=None):
        """Fixtures hit by the segment point1 -> point2 (World.ray_cast, world.py:386-440): each hit as
        [fixture, point, normal, fraction].  As in the reference, the result is a generator: without a
        callback it yields each hit's list, whose last entry the caller may set; with one, iterating
        it calls the callback with each hit and yields nothing.  That value (or the callback's return
        value) steers the cast as in the reference: -1 ignores the hit, 0 terminates, a fraction f
        clips the ray so that hits beyond f are not reported, 1 continues.  Hits are reported in order
        of increasing fraction (ties: fixture order); shapes that contain point1 are not hit.
        Raises ValueError when point1 == point2."""
        p1, p2 = Vec2(*point1), Vec2(*point2)
        if (p2 - p1).length_squared <= 0.0:
            raise ValueError("||p2 - p1||^2 > 0.0")
        seg = np.array([[p1.x, p1.y, p2.x, p2.y]], np.float64)
        rows, data = self._device_state().ray_cast(np.zeros(1, np.int32), seg, closest=False)
        fixtures = self._fixture_list()
        hits = [(fixtures[int(f)], Vec2(d[1], d[2]), Vec2(d[3], d[4]), float(d[0])) for (_, f), d in zip(rows, data)]
        return self._ray_hits(hits, callback)

    @staticmethod
    def _ray_hits(hits, callback):
        max_fraction = 1.0
        for fixture, point, normal, fraction in hits:
            if fraction > max_fraction:
                break
            ret = [fixture, point, normal, fraction]
            if callback is not None:
                value = callback(*ret)
            else:
                yield ret
                value = ret[-1]
            if value == 0.0:
                return
            if value > 0.0:  # DynamicTree.ray_cast: the returned value becomes the max fraction
                max_fraction = value

    def clear_forces(self):
        self._pull()
        for b in self.bodies:
            b._force = Vec2()
            b._torque = 0.0
        self._host_dirty = True

    @property
    def contacts(self) -> List[ContactView]:
        self._pull()
        return [ContactView(self, k, fa, fb) for k, (fa, fb) in enumerate(self._con_fix)]

    @property
    def contact_count(self):
        self._pull()
        return len(self._con_fix)

    def state(self) -> L.BatchArrays:
        """The full world state in the batch layout (after pulling from the device)."""
        self._pull()
        return self._batch()

    @staticmethod
    def default_contact_filter(filter_a, filter_b):
        if filter_a._group_index == filter_b._group_index and filter_a._group_index != 0:
            return filter_a._group_index > 0
        return ((filter_a._mask_bits & filter_b._category_bits) != 0
                and (filter_a._category_bits & filter_b._mask_bits) != 0)


def _transform(d):
    from .common import Transform, Mat22
    t = Transform((d[L.BD_XF_PX], d[L.BD_XF_PY]))
    t._rotation = Mat22((d[L.BD_XF_C], d[L.BD_XF_S]), (-d[L.BD_XF_S], d[L.BD_XF_C]))
    return t
