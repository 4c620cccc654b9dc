"""Joints of the drop-in API (host side): pypybox2d/joints/ mirrored class by class.

Each joint keeps its numbers in one record of the batch layout (b2g_layout.h
B2G_JI_* / B2G_JD_*): its parameters, its accumulated impulses and the values
its solver computed in the last step.  The record crosses the C-ABI with the
world; the device solves the joints inside the islands (island.py:107-148)
and the record comes back with the world's state.  Reading a joint pulls the
device state first; changing one marks the world for upload.

Constructors, properties, setters (which wake the bodies as the reference's
do) and reaction getters follow the reference:
  joint.py:39-112          Joint (same-world check, other_body, collide_connected)
  revolute.py:55-212       RevoluteJoint
  prismatic.py:37-269      PrismaticJoint
  distance.py:58-168       DistanceJoint
  pulley.py:38-179         PulleyJoint
  mouse.py:38-148          MouseJoint (body_a is None)
  gear.py:38-260           GearJoint
  wheel.py:38-239          WheelJoint
  weld.py:38-125           WeldJoint
  friction.py:38-113       FrictionJoint
  rope.py:38-102           RopeJoint
"""
from __future__ import annotations

import numpy as np

from . import layout as L
from .common import EPSILON, Vec2, is_valid_float, scalar_cross
from .settings import LINEAR_SLOP

__all__ = ("Joint", "RevoluteJoint", "PrismaticJoint", "DistanceJoint", "PulleyJoint", "MouseJoint", "GearJoint",
           "WheelJoint", "WeldJoint", "FrictionJoint", "RopeJoint", "INACTIVE_LIMIT", "AT_LOWER_LIMIT",
           "AT_UPPER_LIMIT", "EQUAL_LIMITS", "ALLOWED_STRETCH")

INACTIVE_LIMIT, AT_LOWER_LIMIT, AT_UPPER_LIMIT, EQUAL_LIMITS = range(4)
ALLOWED_STRETCH = 10.0 * LINEAR_SLOP

_PARAMS = 32  # record fields below this are parameters; the rest is solver state


class Joint(object):
    """The base joint class (joint.py:39-112)."""
    TYPE = 0

    def __init__(self, body_a, body_b, collide_connected):
        if body_a and body_b:
            if body_a._world != body_b.world:
                raise ValueError("Both bodies must be in the same world.")
            if body_a._world is None or body_b._world is None:
                raise ValueError("Both bodies must be in the same world.")
        self._body_a = body_a
        self._body_b = body_b
        self._world = None
        self._d = np.zeros(L.JOINT_ND)
        self._i = np.zeros(L.JOINT_NI, np.int32)
        self._i[L.JI_TYPE] = self.TYPE
        self._flag(L.JF_COLLIDE_CONNECTED, collide_connected)

    # ---- record access (pulls the device state first; writes mark the world for upload) ----
    def _pull(self):
        if self._world is not None:
            self._world._pull()

    def _changed(self):
        if self._world is not None:
            self._world._host_dirty = True

    def _get(self, k):
        self._pull()
        return float(self._d[k])

    def _get2(self, k):
        self._pull()
        return Vec2(float(self._d[k]), float(self._d[k + 1]))

    def _put(self, k, v):
        self._pull()
        self._d[k] = v
        self._changed()

    def _flag(self, f, on):
        if on:
            self._i[L.JI_FLAGS] |= f
        else:
            self._i[L.JI_FLAGS] &= ~f

    def _has(self, f):
        return bool(self._i[L.JI_FLAGS] & f)

    def _set_flag(self, f, on):
        self._pull()
        self._flag(f, on)
        self._changed()

    def _wake(self):
        for body in self.bodies:
            body.awake = True

    def _clear_limit_impulse(self):
        # self._impulse -= (0.0, 0.0, self._impulse.z)
        self._d[L.JD_IMP_Z] = self._d[L.JD_IMP_Z] - self._d[L.JD_IMP_Z]

    def _record(self, bidx):
        """The joint's record with its bodies' indices in the world's body list."""
        self._i[L.JI_BODY_A] = bidx[id(self._body_a)] if self._body_a is not None else -1
        self._i[L.JI_BODY_B] = bidx[id(self._body_b)]
        return self._d, self._i

    def __copy__(self):
        """A new joint with the same parameters and no accumulated state (World.add_joint with
        duplicate=True)."""
        j = object.__new__(type(self))
        j.__dict__.update(self.__dict__)
        j._world = None
        j._d = np.zeros(L.JOINT_ND)
        j._d[:_PARAMS] = self._d[:_PARAMS]
        for k in (L.JD_IMP_X, L.JD_IMP_Y, L.JD_IMP_Z, L.JD_MOTOR_IMP, L.JD_SPRING_IMP):
            j._d[k] = 0.0
        j._i = self._i.copy()
        j._i[L.JI_STATE] = INACTIVE_LIMIT
        return j

    # ---- joint.py:57-112 ----
    @property
    def body_a(self):
        """Get the first body attached to this joint."""
        return self._body_a

    @property
    def body_b(self):
        """Get the second body attached to this joint."""
        return self._body_b

    @property
    def bodies(self):
        """Get both bodies attached to this joint"""
        return [self._body_a, self._body_b]

    def other_body(self, body):
        """Given body_a, return body_b or vice-versa."""
        if body == self._body_a:
            return self._body_b
        return self._body_a

    @property
    def active(self):
        """Short-cut function to determine if either body is inactive."""
        return self._body_a.active and self._body_b.active

    @property
    def collide_connected(self):
        """Get collide connected (checked only when fixture AABBs begin to overlap)."""
        return self._has(L.JF_COLLIDE_CONNECTED)

    def get_reaction_force(self, inv_dt):
        """Get the reaction force on body_b at the joint anchor in Newtons."""
        return inv_dt * self._get2(L.JD_IMP_X)

    def get_reaction_torque(self, inv_dt):
        """Get the reaction torque on body_b in N*m."""
        return inv_dt * self._get(L.JD_IMP_Z)

    def _anchors(self, local_anchor_a, local_anchor_b, anchor_a, anchor_b):
        la = Vec2(*local_anchor_a) if local_anchor_a is not None else self._body_a.get_local_point(anchor_a)
        lb = Vec2(*local_anchor_b) if local_anchor_b is not None else self._body_b.get_local_point(anchor_b)
        self._d[L.JD_LA_X], self._d[L.JD_LA_Y] = la.x, la.y
        self._d[L.JD_LB_X], self._d[L.JD_LB_Y] = lb.x, lb.y

    @property
    def anchor_a(self):
        """The world anchor point relative to body_a's origin."""
        return self._body_a.get_world_point(self._get2(L.JD_LA_X))

    @property
    def anchor_b(self):
        """The world anchor point relative to body_b's origin."""
        return self._body_b.get_world_point(self._get2(L.JD_LB_X))


class _LocalAnchors(object):
    @property
    def local_anchor_a(self):
        """The local anchor point relative to body_a's origin."""
        return self._get2(L.JD_LA_X)

    @property
    def local_anchor_b(self):
        """The local anchor point relative to body_b's origin."""
        return self._get2(L.JD_LB_X)


class RevoluteJoint(_LocalAnchors, Joint):
    """Two bodies sharing a point, free to rotate about it; optional angle limits and motor
    (revolute.py:32-453)."""
    TYPE = L.JT_REVOLUTE

    def __init__(self, body_a, body_b, anchor=(0, 0), reference_angle=None, lower_angle=0.0, upper_angle=0.0,
                 max_motor_torque=0.0, motor_speed=0.0, limit_enabled=False, motor_enabled=False,
                 local_anchor_a=None, local_anchor_b=None, collide_connected=False):
        if body_a is None or body_b is None:
            raise ValueError("body_a and body_b must be set")
        Joint.__init__(self, body_a, body_b, collide_connected)
        self._anchors(local_anchor_a, local_anchor_b, anchor, anchor)
        d = self._d
        d[L.JD_REF] = float(reference_angle) if reference_angle is not None else body_b.angle - body_a.angle
        d[L.JD_LOWER], d[L.JD_UPPER] = float(lower_angle), float(upper_angle)
        d[L.JD_MAX_MOTOR] = float(max_motor_torque)
        d[L.JD_MOTOR_SPEED] = float(motor_speed)
        self._flag(L.JF_LIMIT, bool(limit_enabled))
        self._flag(L.JF_MOTOR, bool(motor_enabled))
        for k in range(3):  # _mass = Mat33(): the identity
            d[L.JD_REV_MASS + 4 * k] = 1.0

    @property
    def joint_angle(self):
        """The current joint angle in radians."""
        self._pull()
        return self._body_b._sweep.a - self._body_a._sweep.a - float(self._d[L.JD_REF])

    @property
    def joint_speed(self):
        """The current joint speed in radians/sec."""
        self._pull()
        return self._body_b._angular_velocity - self._body_a._angular_velocity

    @property
    def motor_speed(self):
        """The current motor speed in radians/sec."""
        return self._get(L.JD_MOTOR_SPEED)

    @motor_speed.setter
    def motor_speed(self, speed):
        self._wake()
        self._put(L.JD_MOTOR_SPEED, speed)

    @property
    def limit_enabled(self):
        """The joint limit (min/max angle) enabled"""
        return self._has(L.JF_LIMIT)

    @limit_enabled.setter
    def limit_enabled(self, limit_enabled):
        if limit_enabled != self.limit_enabled:
            self._wake()
            self._set_flag(L.JF_LIMIT, limit_enabled)
            self._d[L.JD_IMP_Z] = 0.0

    @property
    def motor_enabled(self):
        """Joint motor enabled"""
        return self._has(L.JF_MOTOR)

    @motor_enabled.setter
    def motor_enabled(self, motor_enabled):
        if motor_enabled != self.motor_enabled:
            self._set_flag(L.JF_MOTOR, motor_enabled)
            self._wake()

    @property
    def lower_angle(self):
        """The lower angular limit"""
        return self._get(L.JD_LOWER)

    @lower_angle.setter
    def lower_angle(self, lower_angle):
        if self.lower_angle != lower_angle:
            self._wake()
            self._d[L.JD_IMP_Z] = 0.0
            self._put(L.JD_LOWER, lower_angle)

    @property
    def upper_angle(self):
        """The upper angular limit"""
        return self._get(L.JD_UPPER)

    @upper_angle.setter
    def upper_angle(self, upper_angle):
        if self.upper_angle != upper_angle:
            self._wake()
            self._d[L.JD_IMP_Z] = 0.0
            self._put(L.JD_UPPER, upper_angle)

    @property
    def max_motor_torque(self):
        """The maximum motor torque, usually in N-m"""
        return self._get(L.JD_MAX_MOTOR)

    @max_motor_torque.setter
    def max_motor_torque(self, torque):
        self._wake()
        self._put(L.JD_MAX_MOTOR, torque)

    def get_motor_torque(self, inv_dt):
        return inv_dt * self._get(L.JD_MOTOR_IMP)


class PrismaticJoint(Joint):
    """One degree of freedom: translation along an axis fixed in body_a; optional limits and
    motor (prismatic.py:32-513)."""
    TYPE = L.JT_PRISMATIC

    def __init__(self, body_a, body_b, anchor=(0, 0), axis=(0, 1), lower_limit=0.0, upper_limit=0.0,
                 max_motor_force=0.0, max_motor_torque=0.0, motor_speed=0.0, reference_angle=None,
                 limit_enabled=False, motor_enabled=False, collide_connected=False, local_anchor_a=None,
                 local_anchor_b=None, local_x_axis=None):
        if body_a is None or body_b is None:
            raise ValueError("body_a and body_b must be set")
        Joint.__init__(self, body_a, body_b, collide_connected)
        self._anchors(local_anchor_a, local_anchor_b, anchor, anchor)
        d = self._d
        d[L.JD_REF] = float(reference_angle) if reference_angle is not None else body_b.angle - body_a.angle
        lx = Vec2(*local_x_axis) if local_x_axis is not None else body_a.get_local_vector(axis)
        ly = scalar_cross(1.0, lx)
        d[L.JD_AXIS_X], d[L.JD_AXIS_Y] = lx.x, lx.y
        d[L.JD_PERP_X], d[L.JD_PERP_Y] = ly.x, ly.y
        d[L.JD_LOWER], d[L.JD_UPPER] = float(lower_limit), float(upper_limit)
        d[L.JD_MAX_MOTOR] = float(max_motor_force)
        d[L.JD_MOTOR_SPEED] = float(motor_speed)
        self._flag(L.JF_LIMIT, bool(limit_enabled))
        self._flag(L.JF_MOTOR, bool(motor_enabled))

    def get_reaction_force(self, inv_dt):
        """Get the reaction force on body_b at the joint anchor in Newtons."""
        self._pull()
        d = self._d
        perp = Vec2(d[L.JD_PRI_PERP_X], d[L.JD_PRI_PERP_Y])
        axis = Vec2(d[L.JD_PRI_AXIS_X], d[L.JD_PRI_AXIS_Y])
        return inv_dt * (float(d[L.JD_IMP_X]) * perp + (float(d[L.JD_MOTOR_IMP]) + float(d[L.JD_IMP_Z])) * axis)

    def get_reaction_torque(self, inv_dt):
        """Get the reaction torque on body_b in N*m."""
        return inv_dt * self._get(L.JD_IMP_Y)

    @property
    def axis(self):
        return self._body_a.get_world_vector(self._get2(L.JD_AXIS_X))

    @property
    def joint_translation(self):
        """The current joint translation, usually in meters."""
        d = self.anchor_b - self.anchor_a
        return d.dot(self.axis)

    @property
    def joint_speed(self):
        """The current joint translation speed (prismatic.py:145-172).  The reference reads the
        local centres its solver last stored (zero before the first step)."""
        from .common import Mat22
        self._pull()
        ba, bb = self._body_a, self._body_b
        ca, cb = ba.world_center, bb.world_center
        qa, qb = Mat22(angle=ba.angle), Mat22(angle=bb.angle)
        lca = ba._sweep.local_center if self._stepped() else Vec2()
        lcb = bb._sweep.local_center if self._stepped() else Vec2()
        ra = qa * (self._get2(L.JD_LA_X) - lca)
        rb = qb * (self._get2(L.JD_LB_X) - lcb)
        d = (cb + rb) - (ca + ra)
        axis = ba.get_world_vector(self._get2(L.JD_AXIS_X))
        va, vb = ba._linear_velocity, bb._linear_velocity
        wa, wb = ba._angular_velocity, bb._angular_velocity
        return d.dot(scalar_cross(wa, axis)) + axis.dot(vb + scalar_cross(wb, rb) - va - scalar_cross(wa, ra))

    def _stepped(self):
        return bool(np.any(self._d[L.JD_PRI_AXIS_X:L.JD_PRI_PERP_Y + 1] != 0.0))

    @property
    def limit_enabled(self):
        """The joint limit (min/max angle) enabled"""
        return self._has(L.JF_LIMIT)

    @limit_enabled.setter
    def limit_enabled(self, limit_enabled):
        if limit_enabled != self.limit_enabled:
            self._wake()
            self._set_flag(L.JF_LIMIT, limit_enabled)
            self._clear_limit_impulse()

    @property
    def lower_limit(self):
        """The lower translational limit (in meters)"""
        return self._get(L.JD_LOWER)

    @lower_limit.setter
    def lower_limit(self, lower_limit):
        if self.lower_limit != lower_limit:
            self._wake()
            self._clear_limit_impulse()
            self._put(L.JD_LOWER, lower_limit)

    @property
    def upper_limit(self):
        """The upper translational limit (in meters)"""
        return self._get(L.JD_UPPER)

    @upper_limit.setter
    def upper_limit(self, upper_limit):
        if self.upper_limit != upper_limit:
            self._wake()
            self._clear_limit_impulse()
            self._put(L.JD_UPPER, upper_limit)

    @property
    def motor_enabled(self):
        """Joint motor enabled"""
        return self._has(L.JF_MOTOR)

    @motor_enabled.setter
    def motor_enabled(self, motor_enabled):
        if motor_enabled != self.motor_enabled:
            self._set_flag(L.JF_MOTOR, motor_enabled)
            self._wake()

    @property
    def motor_speed(self):
        """The current motor speed, usually in m/s."""
        return self._get(L.JD_MOTOR_SPEED)

    @motor_speed.setter
    def motor_speed(self, speed):
        self._wake()
        self._put(L.JD_MOTOR_SPEED, speed)

    @property
    def max_motor_force(self):
        """The maximum motor force, usually in N"""
        return self._get(L.JD_MAX_MOTOR)

    @max_motor_force.setter
    def max_motor_force(self, force):
        self._wake()
        self._put(L.JD_MAX_MOTOR, force)


class DistanceJoint(_LocalAnchors, Joint):
    """Keeps two anchor points at a fixed distance, or as a spring (distance.py:30-311)."""
    TYPE = L.JT_DISTANCE

    def __init__(self, body_a, body_b, anchor_a=(0, 0), anchor_b=(0, 0), frequency=0.0, damping_ratio=0.0,
                 length=None, collide_connected=False, local_anchor_a=None, local_anchor_b=None):
        if body_a is None or body_b is None:
            raise ValueError("body_a and body_b must be set")
        Joint.__init__(self, body_a, body_b, collide_connected)
        self._anchors(local_anchor_a, local_anchor_b, anchor_a, anchor_b)
        if length is not None:
            length = float(length)
        else:
            length = (self.anchor_b - self.anchor_a).length
        if length <= LINEAR_SLOP:
            raise ValueError("length must be > 0 and not a very small number")
        self._d[L.JD_LENGTH] = length
        self._d[L.JD_FREQ] = frequency
        self._d[L.JD_DAMP] = damping_ratio

    @property
    def length(self):
        """The natural length between the anchor points."""
        return self._get(L.JD_LENGTH)

    @length.setter
    def length(self, length):
        self._put(L.JD_LENGTH, length)

    @property
    def frequency(self):
        """The mass-spring-damper frequency in Hertz."""
        return self._get(L.JD_FREQ)

    @frequency.setter
    def frequency(self, frequency):
        self._put(L.JD_FREQ, frequency)

    @property
    def damping_ratio(self):
        """The damping ratio. 0 = no damping, 1 = critical damping."""
        return self._get(L.JD_DAMP)

    @damping_ratio.setter
    def damping_ratio(self, damping_ratio):
        self._put(L.JD_DAMP, damping_ratio)

    def get_reaction_force(self, inv_dt):
        """Get the reaction force on body_b at the joint anchor in Newtons."""
        return (inv_dt * self._get(L.JD_IMP_X)) * self._get2(L.JD_DIS_U_X)

    def get_reaction_torque(self, inv_dt):
        return 0.0


class PulleyJoint(Joint):
    """Two bodies hung from two ground points by a rope of constant total length over a
    pulley ratio (pulley.py:32-320)."""
    TYPE = L.JT_PULLEY

    def __init__(self, body_a, body_b, ground_anchor_a=(0, 0), ground_anchor_b=(0, 0), anchor_a=(0, 0),
                 anchor_b=(0, 0), ratio=1.0, length_a=None, length_b=None, collide_connected=False,
                 local_anchor_a=None, local_anchor_b=None):
        if body_a is None or body_b is None:
            raise ValueError("body_a and body_b must be set")
        Joint.__init__(self, body_a, body_b, collide_connected)
        if ratio == 0.0 or ratio < EPSILON:
            raise ValueError("Ratio must be positive, non-zero and not too small")
        self._anchors(local_anchor_a, local_anchor_b, anchor_a, anchor_b)
        ga, gb = Vec2(*ground_anchor_a), Vec2(*ground_anchor_b)
        d = self._d
        d[L.JD_G0_X], d[L.JD_G0_Y], d[L.JD_G1_X], d[L.JD_G1_Y] = ga.x, ga.y, gb.x, gb.y
        if length_a is None:
            length_a = (anchor_a - ga).length
        if length_b is None:
            length_b = (anchor_b - gb).length
        d[L.JD_LEN_A], d[L.JD_LEN_B] = length_a, length_b
        d[L.JD_LENGTH] = length_a + ratio * length_b
        d[L.JD_RATIO] = ratio

    def get_reaction_force(self, inv_dt):
        """The reference multiplies by a direction it never sets (pulley.py:92-93): zero."""
        return inv_dt * self._get(L.JD_IMP_X) * Vec2()

    def get_reaction_torque(self, inv_dt):
        return 0.0

    @property
    def length_a(self):
        return (self.anchor_a - self._get2(L.JD_G0_X)).length

    @property
    def length_b(self):
        return (self.anchor_b - self._get2(L.JD_G1_X)).length

    @property
    def ground_anchor_a(self):
        return self._get2(L.JD_G0_X)

    @property
    def ground_anchor_b(self):
        return self._get2(L.JD_G1_X)

    @property
    def ratio(self):
        return self._get(L.JD_RATIO)


class MouseJoint(Joint):
    """Drags a point of one body toward a target with a soft, force-limited constraint
    (mouse.py:32-229).  Its body_a is None."""
    TYPE = L.JT_MOUSE

    def __init__(self, body, target=(0, 0), max_force=0.0, frequency=5.0, damping_ratio=0.7):
        if body is None:
            raise ValueError("body must be set")
        target = Vec2(*target)
        if not target.valid:
            raise ValueError("Invalid target")
        if not is_valid_float(max_force) or max_force < 0.0:
            raise ValueError("Invalid maximum force")
        if not is_valid_float(frequency) or frequency < 0.0:
            raise ValueError("Invalid frequency")
        if not is_valid_float(damping_ratio) or damping_ratio < 0.0:
            raise ValueError("Invalid damping ratio")
        Joint.__init__(self, None, body, False)
        lb = body.get_local_point(target)
        d = self._d
        d[L.JD_G0_X], d[L.JD_G0_Y] = target.x, target.y
        d[L.JD_LB_X], d[L.JD_LB_Y] = lb.x, lb.y
        d[L.JD_MAX_MOTOR] = max_force
        d[L.JD_FREQ] = frequency
        d[L.JD_DAMP] = damping_ratio

    def get_reaction_force(self, inv_dt):
        return inv_dt * self._get2(L.JD_IMP_X)

    def get_reaction_torque(self, inv_dt):
        return 0.0

    @property
    def target(self):
        return self._get2(L.JD_G0_X)

    @target.setter
    def target(self, target):
        if not self._body_b.awake:
            self._body_b.awake = True
        target = Vec2(*target)
        self._put(L.JD_G0_X, target.x)
        self._put(L.JD_G0_Y, target.y)

    @property
    def max_force(self):
        return self._get(L.JD_MAX_MOTOR)

    @max_force.setter
    def max_force(self, max_force):
        self._put(L.JD_MAX_MOTOR, max_force)

    @property
    def frequency(self):
        return self._get(L.JD_FREQ)

    @frequency.setter
    def frequency(self, frequency):
        self._put(L.JD_FREQ, frequency)

    @property
    def damping_ratio(self):
        return self._get(L.JD_DAMP)

    @damping_ratio.setter
    def damping_ratio(self, damping_ratio):
        self._put(L.JD_DAMP, damping_ratio)


class GearJoint(Joint):
    """Couples two revolute or prismatic joints (gear.py:32-388).  As in the reference, its
    bodies a and b are both joint_a's body_b, and c and d both joint_a's body_a (gear.py:71-92)."""
    TYPE = L.JT_GEAR

    def __init__(self, joint_a, joint_b, ratio=0.1):
        Joint.__init__(self, None, None, False)
        if not isinstance(joint_a, (RevoluteJoint, PrismaticJoint)):
            raise ValueError("Joints must be either Revolute or Prismatic")
        if not isinstance(joint_b, (RevoluteJoint, PrismaticJoint)):
            raise ValueError("Joints must be either Revolute or Prismatic")
        self._joint_a, self._joint_b = joint_a, joint_b
        joint_a._pull()
        d = self._d
        body_a = self._body_a = joint_a.body_b
        body_c = self._body_c = joint_a.body_a
        xf_a, xf_c = body_a._xf, body_c._xf
        aa, ac = body_a._sweep.a, body_c._sweep.a
        self._revolute_a = isinstance(joint_a, RevoluteJoint)
        ja, jb = joint_a._d, joint_b._d
        lc, la = Vec2(ja[L.JD_LA_X], ja[L.JD_LA_Y]), Vec2(ja[L.JD_LB_X], ja[L.JD_LB_Y])
        d[L.JD_G0_X], d[L.JD_G0_Y] = lc.x, lc.y
        d[L.JD_LA_X], d[L.JD_LA_Y] = la.x, la.y
        d[L.JD_REF] = ja[L.JD_REF]
        if self._revolute_a:
            coordinate_a = aa - ac - float(ja[L.JD_REF])
        else:
            axis_c = Vec2(ja[L.JD_AXIS_X], ja[L.JD_AXIS_Y])
            d[L.JD_AXIS_X], d[L.JD_AXIS_Y] = axis_c.x, axis_c.y
            pa = xf_c.rotation.mul_t(xf_a.rotation * la + (xf_a.position - xf_c.position))
            coordinate_a = (pa - lc).dot(axis_c)
        body_d = self._body_d = joint_a.body_a
        body_b = self._body_b = joint_a.body_b
        xf_b, xf_d = body_b._xf, body_d._xf
        ab, ad = body_b._sweep.a, body_d._sweep.a
        self._revolute_b = isinstance(joint_b, RevoluteJoint)
        ld, lb = Vec2(jb[L.JD_LA_X], jb[L.JD_LA_Y]), Vec2(jb[L.JD_LB_X], jb[L.JD_LB_Y])
        d[L.JD_G1_X], d[L.JD_G1_Y] = ld.x, ld.y
        d[L.JD_LB_X], d[L.JD_LB_Y] = lb.x, lb.y
        d[L.JD_REF_B] = jb[L.JD_REF]
        if self._revolute_b:
            coordinate_b = ab - ad - float(jb[L.JD_REF])
        else:
            axis_d = Vec2(jb[L.JD_AXIS_X], jb[L.JD_AXIS_Y])
            d[L.JD_PERP_X], d[L.JD_PERP_Y] = axis_d.x, axis_d.y
            pb = xf_d.rotation.mul_t(xf_b.rotation * lb + (xf_b.position - xf_d.position))
            coordinate_b = (pb - ld).dot(axis_d)
        self._flag(L.JF_REVOLUTE_A, self._revolute_a)
        self._flag(L.JF_REVOLUTE_B, self._revolute_b)
        d[L.JD_RATIO] = ratio
        d[L.JD_LENGTH] = coordinate_a + ratio * coordinate_b

    def _record(self, bidx):
        Joint._record(self, bidx)
        self._i[L.JI_BODY_C] = bidx[id(self._body_c)]
        self._i[L.JI_BODY_D] = bidx[id(self._body_d)]
        return self._d, self._i

    def get_reaction_force(self, inv_dt):
        raise AttributeError("'GearJoint' object has no attribute '_j' (gear.py:166)")

    def get_reaction_torque(self, inv_dt):
        raise AttributeError("'GearJoint' object has no attribute '_j' (gear.py:171)")

    @property
    def joint_a(self):
        return self._joint_a

    @property
    def joint_b(self):
        return self._joint_b

    @property
    def revolute_joint_a(self):
        return self._revolute_a

    @property
    def revolute_joint_b(self):
        return self._revolute_b

    @property
    def ratio(self):
        return self._get(L.JD_RATIO)

    @ratio.setter
    def ratio(self, ratio):
        if not is_valid_float(ratio):
            raise ValueError("Invalid ratio")
        self._put(L.JD_RATIO, ratio)


class WheelJoint(_LocalAnchors, Joint):
    """A point of body_b on a line of body_a, with a spring along the line's normal and a
    rotation motor (wheel.py:32-417)."""
    TYPE = L.JT_WHEEL

    def __init__(self, body_a, body_b, anchor=None, axis=None, frequency=2.0, damping_ratio=0.7,
                 motor_enabled=False, motor_speed=0.0, max_motor_torque=0.0, local_anchor_a=None,
                 local_anchor_b=None, local_x_axis=None, collide_connected=False):
        if body_a is None or body_b is None:
            raise ValueError("body_a and body_b must be set")
        Joint.__init__(self, body_a, body_b, collide_connected)
        self._anchors(local_anchor_a, local_anchor_b, anchor, anchor)
        d = self._d
        lx = Vec2(*local_x_axis) if local_x_axis is not None else body_a.get_local_vector(axis)
        ly = scalar_cross(1.0, lx)
        d[L.JD_AXIS_X], d[L.JD_AXIS_Y] = lx.x, lx.y
        d[L.JD_PERP_X], d[L.JD_PERP_Y] = ly.x, ly.y
        d[L.JD_FREQ], d[L.JD_DAMP] = float(frequency), float(damping_ratio)
        d[L.JD_MAX_MOTOR], d[L.JD_MOTOR_SPEED] = float(max_motor_torque), float(motor_speed)
        self._flag(L.JF_MOTOR, bool(motor_enabled))

    def get_reaction_force(self, inv_dt):
        self._pull()
        d = self._d
        ay, ax = Vec2(d[L.JD_WHE_AY_X], d[L.JD_WHE_AY_Y]), Vec2(d[L.JD_WHE_AX_X], d[L.JD_WHE_AX_Y])
        return inv_dt * (float(d[L.JD_IMP_X]) * ay + float(d[L.JD_SPRING_IMP]) * ax)

    def get_reaction_torque(self, inv_dt):
        return inv_dt * self._get(L.JD_MOTOR_IMP)

    @property
    def joint_translation(self):
        d = self.anchor_b - self.anchor_a
        axis = self._body_a.get_world_vector(self._get2(L.JD_AXIS_X))
        return d.dot(axis)

    @property
    def joint_speed(self):
        self._pull()
        return self._body_b._angular_velocity - self._body_a._angular_velocity

    @property
    def motor_enabled(self):
        return self._has(L.JF_MOTOR)

    @motor_enabled.setter
    def motor_enabled(self, motor_enabled):
        if motor_enabled != self.motor_enabled:
            self._wake()
            self._set_flag(L.JF_MOTOR, motor_enabled)

    @property
    def motor_speed(self):
        return self._get(L.JD_MOTOR_SPEED)

    @motor_speed.setter
    def motor_speed(self, motor_speed):
        self._wake()
        self._put(L.JD_MOTOR_SPEED, motor_speed)

    @property
    def max_motor_torque(self):
        return self._get(L.JD_MAX_MOTOR)

    @max_motor_torque.setter
    def max_motor_torque(self, max_motor_torque):
        self._wake()
        self._put(L.JD_MAX_MOTOR, max_motor_torque)

    def motor_torque(self, inv_dt):
        return inv_dt * self._get(L.JD_MOTOR_IMP)

    # The reference's spring setters store attributes the solver never reads (wheel.py:221-239):
    # the spring keeps the constructor's frequency and damping ratio.
    @property
    def spring_frequency(self):
        return self._spring_frequency

    @spring_frequency.setter
    def spring_frequency(self, spring_frequency):
        self._spring_frequency = spring_frequency

    @property
    def spring_damping_ratio(self):
        return self._spring_damping_ratio

    @spring_damping_ratio.setter
    def spring_damping_ratio(self, spring_damping_ratio):
        self._spring_damping_ratio = spring_damping_ratio


class WeldJoint(Joint):
    """Glues two bodies together at an anchor (weld.py:32-257)."""
    TYPE = L.JT_WELD

    def __init__(self, body_a, body_b, anchor=None, reference_angle=None, local_anchor_a=None, local_anchor_b=None,
                 collide_connected=False):
        if body_a is None or body_b is None:
            raise ValueError("body_a and body_b must be set")
        Joint.__init__(self, body_a, body_b, collide_connected)
        if anchor is not None:
            self._anchors(None, None, anchor, anchor)
        else:
            self._anchors(local_anchor_a, local_anchor_b, None, None)
        d = self._d
        d[L.JD_REF] = float(reference_angle) if reference_angle is not None else body_b.angle - body_a.angle
        for k in range(3):  # _mass = Mat33(): the identity
            d[L.JD_WEL_MASS + 4 * k] = 1.0


class FrictionJoint(Joint):
    """Top-down friction: resists relative motion up to a maximum force and torque
    (friction.py:32-233)."""
    TYPE = L.JT_FRICTION

    def __init__(self, body_a, body_b, local_anchor_a=(0, 0), local_anchor_b=(0, 0), max_force=0.0, max_torque=0.0,
                 collide_connected=False):
        if body_a is None or body_b is None:
            raise ValueError("body_a and body_b must be set")
        Joint.__init__(self, body_a, body_b, collide_connected)
        self._anchors(local_anchor_a, local_anchor_b, None, None)
        self._d[L.JD_MAX_MOTOR] = max_force
        self._d[L.JD_MAX_TORQUE] = max_torque

    def get_reaction_force(self, inv_dt):
        return inv_dt * self._get2(L.JD_IMP_X)

    def get_reaction_torque(self, inv_dt):
        return inv_dt * self._get(L.JD_IMP_Z)

    @property
    def max_force(self):
        return self._get(L.JD_MAX_MOTOR)

    @max_force.setter
    def max_force(self, max_force):
        if not is_valid_float(max_force) or max_force < 0.0:
            raise ValueError("Max force must be >= 0.0")
        self._put(L.JD_MAX_MOTOR, max_force)

    @property
    def max_torque(self):
        return self._get(L.JD_MAX_TORQUE)

    @max_torque.setter
    def max_torque(self, max_torque):
        if not is_valid_float(max_torque) or max_torque < 0.0:
            raise ValueError("Max torque must be >= 0.0")
        self._put(L.JD_MAX_TORQUE, max_torque)


class RopeJoint(Joint):
    """An upper bound on the distance between two anchor points (rope.py:32-248)."""
    TYPE = L.JT_ROPE

    def __init__(self, body_a, body_b, local_anchor_a=(0, 0), local_anchor_b=(0, 0), max_length=0.0,
                 collide_connected=False):
        if body_a is None or body_b is None:
            raise ValueError("body_a and body_b must be set")
        Joint.__init__(self, body_a, body_b, collide_connected)
        self._anchors(local_anchor_a, local_anchor_b, None, None)
        self._d[L.JD_LENGTH] = max_length

    def get_reaction_force(self, inv_dt):
        return (inv_dt * self._get(L.JD_IMP_X)) * self._get2(L.JD_ROP_U_X)

    def get_reaction_torque(self, inv_dt):
        return 0.0

    @property
    def limit_state(self):
        raise AttributeError("'RopeJoint' object has no attribute '_limit_state' (rope.py:88)")

    @property
    def max_length(self):
        return self._get(L.JD_LENGTH)
