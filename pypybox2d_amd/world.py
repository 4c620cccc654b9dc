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

The contact listener (``world.contact_manager.begin_contact``, ``end_contact``,
``pre_solve``, ``post_solve``) is called with the device's records of the step,
in the reference's call order: the step runs in two parts, and the callbacks
of ContactManager.collide (begin, end, pre_solve) run in between, so that
``contact.enabled = False`` keeps a contact out of the step's solve as in the
reference; post_solve follows the solve.  Inside a callback the world is
locked: creating or destroying bodies raises LockedError as in the reference,
and so do Body setters (the rest of the step is already on the device).  A
custom ``contact_filter`` runs on the host on the contacts each pair search
creates (the device creates every pair Body._should_collide admits; the
rejected ones are dropped before the next step) and on the contacts flagged
by Fixture.refilter, before the step that would test them.

Continuous physics (``continuous_physics=True``, the reference's default) runs
on the device too: after each step the time-of-impact sub-steps of
World.solve_toi (world.py:594-821) run per world.  Its listener calls come
after the step's others.

Outside the device path, and therefore refused with NotImplementedError when a
world holding them is stepped: joints, controllers, sensors, loop shapes,
sub-stepping, and a custom contact filter together with continuous physics.
"""
from __future__ import annotations

from copy import copy
from typing import List, Optional

import numpy as np

from . import joints
from . import layout as L
from .body import Body
from .common import LockedError, Vec2
from .contact import Contact, ContactManager, Manifold
from .settings import AABB_EXTENSION
from .shapes import Circle, Edge, Loop, Polygon, check_device_shape

__all__ = ("World", "TimeStep")


class TimeStep(object):
    __slots__ = ("dt", "inv_dt", "dt_ratio", "vel_iters", "pos_iters", "warm_starting")

    def __init__(self, dt=0.0, inv_dt=0.0, dt_ratio=0.0, vel_iters=0, pos_iters=0, warm_starting=True):
        self.dt, self.inv_dt, self.dt_ratio = float(dt), float(inv_dt), float(dt_ratio)
        self.vel_iters, self.pos_iters, self.warm_starting = int(vel_iters), int(pos_iters), bool(warm_starting)


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
        self.contact_manager = ContactManager(self)
        self._contact_filter = self.default_contact_filter if contact_filter is None else contact_filter
        self._views = {}              # (id(fixture a), id(fixture b)) -> Contact, while the contact exists
        self._refilter_pending = False  # contacts flagged for filtering wait for a custom filter's verdict
        self._deferred_refilter: List = []  # fixtures whose filter bits a listener callback changed
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
        self._pull()
        self._continuous_physics = bool(v)
        self._host_dirty = True

    @property
    def sub_stepping(self):
        return self._sub_stepping

    @sub_stepping.setter
    def sub_stepping(self, v):
        self._sub_stepping = bool(v)

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
        """world.py:124-136.  None turns filtering off (every pair collides)."""
        self._pull()
        self._contact_filter = fcn
        self._refilter_pending = True
        self._host_dirty = True

    def _custom_filter(self) -> bool:
        return self._contact_filter is not World.default_contact_filter

    def _listening(self) -> bool:
        return any(self.contact_manager.contact_listener)

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
        # Body._detached_from_world (body.py:310-345): its joints are destroyed first (the joint
        # destruction listener, then World.destroy_joint)
        for joint in list(body.joints):
            if joint in self.joints:
                if self.joint_destruction_listener is not None:
                    self.joint_destruction_listener(joint)
                self.destroy_joint(joint)
        del body.joints[:]
        fixtures = set(id(f) for f in body.fixtures)
        # then its contacts (end_contact for the touching ones, ContactManager.destroy), then the
        # fixture destruction listener is called
        self._destroy_contacts(lambda fa, fb: id(fa) in fixtures or id(fb) in fixtures)
        if self.fixture_destruction_listener is not None:
            for f in list(body.fixtures):
                self.fixture_destruction_listener(f)
        self._move = [f for f in self._move if id(f) not in fixtures]
        self.bodies.remove(body)
        body._world = None
        self._host_dirty = True

    # ---- joints (world.py:252-306, 823-921) ----
    def add_joint(self, joint, duplicate=False):
        """Create a joint.  Creating a joint doesn't wake the bodies; contacts between its bodies
        are flagged for filtering when the joint keeps them from colliding."""
        if self._locked:
            raise LockedError("Cannot create a joint while simulating")
        self._pull()
        if duplicate:
            joint = copy(joint)
        self.joints.append(joint)
        joint._world = self
        if joint._body_a is not None:
            joint._body_a.joints.append(joint)
        if joint._body_b is not None:
            joint._body_b.joints.append(joint)
            if not joint._has(L.JF_COLLIDE_CONNECTED):
                self._flag_between(joint._body_b, joint._body_a)
        self._host_dirty = True
        return joint

    def destroy_joint(self, joint):
        """Destroy an already created joint; its bodies are woken, and contacts between them
        are flagged for filtering."""
        if self._locked:
            raise LockedError("Cannot destroy a joint while simulating")
        if joint not in self.joints:
            raise ValueError("Joint not in world")
        self._pull()
        self.joints.remove(joint)
        for body in joint.bodies:
            if body is not None:
                body.awake = True
                body.joints.remove(joint)
        if not joint._has(L.JF_COLLIDE_CONNECTED) and joint._body_b is not None:
            self._flag_between(joint._body_b, joint._body_a)
        joint._world = None
        self._host_dirty = True

    def _flag_between(self, body, other):
        """contact._flag_for_filtering for body's contacts whose other body is `other`."""
        for k, (fa, fb) in enumerate(self._con_fix):
            if (fa._body is body and fb._body is other) or (fb._body is body and fa._body is other):
                self._con_i[k, L.CI_FLAGS] |= L.CF_FILTER
                self._refilter_pending = True

    def create_distance_joint(self, *args, **kwargs):
        return self.add_joint(joints.DistanceJoint(*args, **kwargs))

    def create_weld_joint(self, *args, **kwargs):
        return self.add_joint(joints.WeldJoint(*args, **kwargs))

    def create_friction_joint(self, *args, **kwargs):
        return self.add_joint(joints.FrictionJoint(*args, **kwargs))

    def create_gear_joint(self, *args, **kwargs):
        return self.add_joint(joints.GearJoint(*args, **kwargs))

    def create_mouse_joint(self, *args, **kwargs):
        return self.add_joint(joints.MouseJoint(*args, **kwargs))

    def create_prismatic_joint(self, *args, **kwargs):
        return self.add_joint(joints.PrismaticJoint(*args, **kwargs))

    def create_pulley_joint(self, *args, **kwargs):
        return self.add_joint(joints.PulleyJoint(*args, **kwargs))

    def create_revolute_joint(self, *args, **kwargs):
        return self.add_joint(joints.RevoluteJoint(*args, **kwargs))

    def create_rope_joint(self, *args, **kwargs):
        return self.add_joint(joints.RopeJoint(*args, **kwargs))

    def create_wheel_joint(self, *args, **kwargs):
        return self.add_joint(joints.WheelJoint(*args, **kwargs))

    # hooks used by Body / Fixture
    def _fixture_added(self, body, fixture):
        if body._active:
            fixture._fat = fat_aabb(fixture._shape, body._xf)
            self._move.append(fixture)
        self._new_fixture = True
        self._host_dirty = True

    def _fixture_removed(self, body, fixture):
        """Body.destroy_fixture (body.py:258-286): the fixture's contacts are destroyed."""
        self._destroy_contacts(lambda fa, fb: fa is fixture or fb is fixture)
        self._move = [f for f in self._move if f is not fixture]
        self._host_dirty = True

    def _refilter(self, fixture):
        """Fixture.refilter (fixture.py:269-285): the fixture's contacts are flagged for filtering
        at the next step where a body of theirs is awake, and its proxy is touched so that the pair
        search reconsiders its pairs."""
        self._pull()
        for k, (fa, fb) in enumerate(self._con_fix):
            if fa is fixture or fb is fixture:
                self._con_i[k, L.CI_FLAGS] |= L.CF_FILTER
        if fixture._body._active and fixture not in self._move:
            self._move.append(fixture)
        self._refilter_pending = True
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

    def _destroy_contacts(self, pred):
        """ContactManager.destroy (contact_manager.py:160-172) of every contact matching pred, in
        list order: end_contact for those touching, then removal."""
        end = self.contact_manager.end_contact
        gone = [k for k, (fa, fb) in enumerate(self._con_fix) if pred(fa, fb)]
        views = [self._row_view(k) for k in gone]
        self._drop_contacts(pred)
        for c in views:
            self._views.pop((id(c._fixture_a), id(c._fixture_b)), None)
            if end is not None and c._touching_flag:
                end(c)

    def _view(self, fa, fb) -> Contact:
        """The contact object of a fixture pair (the same object while the contact exists)."""
        key = (id(fa), id(fb))
        c = self._views.get(key)
        if c is None:
            c = Contact(self, fa, fb)
            self._views[key] = c
        return c

    def _row_view(self, k) -> Contact:
        fa, fb = self._con_fix[k]
        c = self._view(fa, fb)
        c._load_row(self._con_d[k], self._con_i[k])
        return c

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
                             | (0 if self._auto_clear_forces else L.WF_KEEP_FORCES)
                             | (L.WF_CUSTOM_FILTER if self._custom_filter() else 0)
                             | (L.WF_CONTINUOUS if self._continuous_physics else 0))
        nj = len(self.joints)
        jd = np.zeros((nj, L.JOINT_ND))
        ji = np.zeros((nj, L.JOINT_NI), np.int32)
        for k, j in enumerate(self.joints):
            jd[k], ji[k] = j._record(bidx)
        return L.BatchArrays(np.array([0, nb], np.int32), np.array([0, nf], np.int32),
                             np.array([0, len(ci)], np.int32), np.array([0, len(move)], np.int32),
                             wd, wi, bd, bi, fd, fi, self._con_d.copy(), ci, move,
                             np.array([0, nj], np.int32), jd, ji)

    def _check_supported(self):
        if self._continuous_physics and self._sub_stepping:
            raise NotImplementedError("sub-stepping is not supported")
        if self._continuous_physics and self._custom_filter():
            raise NotImplementedError("a custom contact filter with continuous physics is not supported (the "
                                      "time-of-impact sub-steps create contacts inside the device step)")
        if self.controllers:
            raise NotImplementedError("controllers are outside the device contact pipeline")
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
        for j, d, i in zip(self.joints, st.joint_d, st.joint_i):
            j._d[:] = d
            j._i[:] = i
        self._inv_dt0 = float(st.world_d[0, L.WD_INV_DT0])
        self._new_fixture = bool(st.world_i[0, L.WI_FLAGS] & L.WF_NEW_FIXTURE)

    def step(self, dt, vel_iters, pos_iters):
        """One World.step (world.py:307-354) on the GPU."""
        if self._locked:
            raise LockedError
        custom = self._custom_filter()
        if custom and self._refilter_pending:
            self._pull()
            self._filter_flagged()
        if self._host_dirty or self._stepper is None:
            self._pull()
            self._push()
        self._locked = True
        try:
            # steps, then checks the device's error flags: a capacity error grows the contact
            # slots and replays from the last synchronised state; non-finite state raises
            if custom or self._listening():
                self._step_listened(float(dt), int(vel_iters), int(pos_iters), custom)
            else:
                self._stepper.step(float(dt), int(vel_iters), int(pos_iters))
            self._device_ahead = True
            self._new_fixture = False
            if dt > 0.0:
                self._inv_dt0 = 1.0 / dt
        finally:
            self._locked = False
        deferred, self._deferred_refilter = self._deferred_refilter, []
        for f in deferred:
            if f.world is self:
                self._refilter(f)

    # ---- contact listener and custom filter ----
    def _filter_flagged(self):
        """ContactManager.collide's test of the contacts flagged for filtering (contact_manager.py:
        192-208) with the custom filter, for the flagged contacts the coming step will test (a body
        of theirs awake and not static): Body._should_collide, then the filter.  The device destroys
        the rejected ones in that step's collide (B2G_CF_REJECT); the others keep no flag."""
        pending = False
        for k, (fa, fb) in enumerate(self._con_fix):
            fl = int(self._con_i[k, L.CI_FLAGS])
            if not fl & L.CF_FILTER or fl & L.CF_REJECT:
                continue
            ba, bb = fa._body, fb._body
            if not ((ba._awake and ba._type != Body.STATIC) or (bb._awake and bb._type != Body.STATIC)):
                pending = True
                continue
            ok = bb._should_collide(ba) and (self._contact_filter is None or bool(self._contact_filter(fa, fb)))
            self._con_i[k, L.CI_FLAGS] = (fl & ~L.CF_FILTER) if ok else (fl | L.CF_REJECT)
            self._host_dirty = True
        self._refilter_pending = pending

    def _step_listened(self, dt, vel_iters, pos_iters, custom):
        st = self._stepper
        cm = self.contact_manager
        mask = ((1 << L.EV_BEGIN) | (1 << L.EV_END) | ((1 << L.EV_PRE_SOLVE) if cm.pre_solve else 0)
                | ((1 << L.EV_POST_SOLVE) if cm.post_solve else 0))
        if st.params.get("events") != mask:
            st.set_param("events", mask)
        fixtures = self._fixture_list()
        if custom and self._new_fixture:  # the new fixtures' pairs, filtered before the step uses them
            st.pair_search()
            self._new_contacts(fixtures, custom)
        st.collide(dt, vel_iters, pos_iters)
        ei, ed = st.events()
        disable = self._dispatch(ei, ed, fixtures, (0, 1))
        if disable:
            st.disable([0] * len(disable), disable)
        st.solve()
        self._device_ahead = True
        if cm.post_solve is not None or self._continuous_physics:
            ei, ed = st.events()  # post_solve, then the calls of the time-of-impact sub-steps
            self._dispatch(ei, ed, fixtures, (2, 3))
        self._new_contacts(fixtures, custom)

    def _dispatch(self, ei, ed, fixtures, phases):
        """Calls the listener for the records of the given phases, in order; returns the contact
        slots whose contact a callback disabled."""
        cm = self.contact_manager
        updated = {}
        for r in range(ei.shape[0]):
            row = ei[r]
            phase = int(row[L.EI_PHASE])
            if phase not in phases:
                continue
            c = self._view(fixtures[row[L.EI_FA]], fixtures[row[L.EI_FB]])
            c._load_event(row, ed[r])
            kind = int(row[L.EI_KIND])
            if phase == 0:
                updated[int(row[L.EI_KEY]) >> 2] = c
            if kind == L.EV_BEGIN:
                if cm.begin_contact is not None:
                    cm.begin_contact(c)
            elif kind == L.EV_END:
                if cm.end_contact is not None:
                    cm.end_contact(c)
                if phase == 1:  # destroyed
                    self._views.pop((id(c._fixture_a), id(c._fixture_b)), None)
            elif kind == L.EV_PRE_SOLVE:
                if cm.pre_solve is not None:
                    cm.pre_solve(c, Manifold.from_record(ed[r, L.ED_OLD:], row[L.EI_OLD_MTYPE:L.EI_OLD_MTYPE + 4]))
            elif kind == L.EV_POST_SOLVE:
                if cm.post_solve is not None:
                    o = L.ED_OLD
                    cm.post_solve(c, [(float(ed[r, o + L.ED_P0_NI + 4 * j]), float(ed[r, o + L.ED_P0_TI + 4 * j]))
                                      for j in range(int(row[L.EI_OLD_PC]))])
        return sorted(slot for slot, c in updated.items() if not c._enabled_flag)

    def _new_contacts(self, fixtures, custom):
        """The contacts the last pair search created: each is a new contact object, and with a custom
        filter (contact_manager.py:144-146) each is offered to it as the pair search reported it
        (registry swap undone); the rejected ones are dropped."""
        rows = self._stepper.new_contacts()
        drop = []
        f = self._contact_filter
        for j in range(rows.shape[0]):
            a, b = fixtures[rows[j, 1]], fixtures[rows[j, 2]]
            self._views.pop((id(a), id(b)), None)
            if custom and f is not None:
                p, q = (b, a) if rows[j, 3] & L.CF_SWAPPED else (a, b)
                if not f(p, q):
                    drop.append(j)
        if drop:
            self._stepper.drop([0] * len(drop), drop)

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
    def contacts(self) -> List[Contact]:
        """The contact list (contact_manager.contacts), in the device's list order."""
        self._pull()
        out = [self._row_view(k) for k in range(len(self._con_fix))]
        if not self._locked:  # forget the contacts that no longer exist
            self._views = {(id(c._fixture_a), id(c._fixture_b)): c for c in out}
        return out

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
