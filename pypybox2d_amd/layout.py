"""Record layout of a batch of worlds as it crosses the C-ABI.

Mirrors ``include/b2g_layout.h`` field for field (the header is the source of
truth; ``tests/test_layout.py`` parses it and checks every constant here).
A batch is a set of numpy arrays of fixed-stride float64 / int32 rows plus
per-world offset arrays; ``BatchArrays.as_struct`` exposes them as the
``b2g_batch`` C struct without copying.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

MAX_POLY_VERTS = 8

SHAPE_CIRCLE, SHAPE_POLYGON, SHAPE_LOOP, SHAPE_EDGE = 0, 1, 2, 3
STATIC, DYNAMIC, KINEMATIC = 0, 1, 2
M_CIRCLES, M_FACE_A, M_FACE_B = 0, 1, 2

# body float64 record
(BD_XF_PX, BD_XF_PY, BD_XF_C, BD_XF_S, BD_LC_X, BD_LC_Y, BD_C0_X, BD_C0_Y, BD_C_X, BD_C_Y,
 BD_A0, BD_A, BD_V_X, BD_V_Y, BD_W, BD_F_X, BD_F_Y, BD_TORQUE, BD_MASS, BD_INV_MASS, BD_I,
 BD_INV_I, BD_GRAV_SCALE, BD_LIN_DAMP, BD_ANG_DAMP, BD_SLEEP_TIME) = range(26)
BODY_ND = 28
BI_TYPE, BI_FLAGS, BI_ISLAND, BI_PAD = range(4)
BODY_NI = 4
BF_AWAKE, BF_ALLOW_SLEEP, BF_FIXED_ROT, BF_ACTIVE, BF_BULLET = 1, 2, 4, 8, 16

# fixture float64 record
(FD_FRICTION, FD_RESTITUTION, FD_DENSITY, FD_RADIUS, FD_FAT_LX, FD_FAT_LY, FD_FAT_UX,
 FD_FAT_UY, FD_CENTROID_X, FD_CENTROID_Y, FD_CIRCLE_PX, FD_CIRCLE_PY, FD_EDGE_V0X,
 FD_EDGE_V0Y, FD_EDGE_V1X, FD_EDGE_V1Y, FD_EDGE_V2X, FD_EDGE_V2Y, FD_EDGE_V3X,
 FD_EDGE_V3Y) = range(20)
FD_VERTS = 20
FD_NORMALS = FD_VERTS + 2 * MAX_POLY_VERTS
FIX_ND = 56
(FI_BODY, FI_SHAPE, FI_NVERTS, FI_EDGE_FLAGS, FI_CATEGORY, FI_MASK, FI_GROUP,
 FI_SENSOR) = range(8)
FIX_NI = 8
EF_HAS_V0, EF_HAS_V3 = 1, 2

# contact float64 record
(CD_FRICTION, CD_RESTITUTION, CD_LN_X, CD_LN_Y, CD_LP_X, CD_LP_Y, CD_P0_X, CD_P0_Y, CD_P0_NI,
 CD_P0_TI, CD_P1_X, CD_P1_Y, CD_P1_NI, CD_P1_TI) = range(14)
CON_ND = 16
CI_FA, CI_FB, CI_FLAGS, CI_MTYPE, CI_PC, CI_ID0, CI_ID1, CI_PAD = range(8)
CON_NI = 8
CF_TOUCHING, CF_ENABLED, CF_FILTER = 1, 2, 4
CF_SWAPPED = 8   # b2g_new_contacts rows: the registry swapped the reported pair (B, A)
CF_REJECT = 16   # with WF_CUSTOM_FILTER: the host's filter rejected this flagged contact
NC_DROPPED = 16  # b2g_new_contacts rows: dropped since

WD_GX, WD_GY, WD_INV_DT0 = range(3)
WORLD_ND = 4
WI_FLAGS, WI_NEW_CONTACTS = 0, 1
WORLD_NI = 4
WF_ALLOW_SLEEP, WF_WARM_STARTING, WF_NEW_FIXTURE, WF_KEEP_FORCES, WF_CUSTOM_FILTER = 1, 2, 4, 8, 16
WF_CONTINUOUS = 32  # continuous physics on (World.solve_toi after each step)

# contact listener records (b2g_events)
EV_BEGIN, EV_END, EV_PRE_SOLVE, EV_POST_SOLVE = range(4)
EV_ALL = 15
(EI_WORLD, EI_KIND, EI_PHASE, EI_KEY, EI_FA, EI_FB, EI_FLAGS, EI_MTYPE, EI_PC, EI_ID0, EI_ID1,
 EI_OLD_MTYPE, EI_OLD_PC, EI_OLD_ID0, EI_OLD_ID1, EI_PAD) = range(16)
EV_NI = 16
(ED_LN_X, ED_LN_Y, ED_LP_X, ED_LP_Y, ED_P0_X, ED_P0_Y, ED_P0_NI, ED_P0_TI, ED_P1_X, ED_P1_Y, ED_P1_NI,
 ED_P1_TI) = range(12)
ED_OLD = 12
EV_ND = 24


# joint records (b2g_layout.h: world joints in World.joints order)
JT_REVOLUTE, JT_PRISMATIC, JT_DISTANCE, JT_PULLEY, JT_MOUSE, JT_GEAR, JT_WHEEL, JT_WELD, JT_FRICTION, \
    JT_ROPE = range(1, 11)
JI_TYPE, JI_BODY_A, JI_BODY_B, JI_FLAGS, JI_STATE, JI_BODY_C, JI_BODY_D, JI_PAD = range(8)
JOINT_NI = 8
JF_COLLIDE_CONNECTED, JF_LIMIT, JF_MOTOR, JF_REVOLUTE_A, JF_REVOLUTE_B = 1, 2, 4, 8, 16
LIMIT_INACTIVE, LIMIT_AT_LOWER, LIMIT_AT_UPPER, LIMIT_EQUAL = range(4)
(JD_LA_X, JD_LA_Y, JD_LB_X, JD_LB_Y, JD_REF, JD_IMP_X, JD_IMP_Y, JD_IMP_Z, JD_MOTOR_IMP, JD_MOTOR_SPEED,
 JD_MAX_MOTOR, JD_LOWER, JD_UPPER, JD_AXIS_X, JD_AXIS_Y, JD_PERP_X, JD_PERP_Y, JD_FREQ, JD_DAMP, JD_LENGTH,
 JD_RATIO, JD_G0_X, JD_G0_Y, JD_G1_X, JD_G1_Y, JD_SPRING_IMP, JD_MAX_TORQUE, JD_REF_B, JD_LEN_A,
 JD_LEN_B) = range(30)
JD_RA_X, JD_RA_Y, JD_RB_X, JD_RB_Y = range(32, 36)
JD_S = 36
JD_REV_MASS, JD_REV_MOTOR_MASS = JD_S, JD_S + 9
(JD_PRI_AXIS_X, JD_PRI_AXIS_Y, JD_PRI_PERP_X, JD_PRI_PERP_Y, JD_PRI_A1, JD_PRI_A2, JD_PRI_S1, JD_PRI_S2,
 JD_PRI_MOTOR_MASS, JD_PRI_K) = range(JD_S, JD_S + 10)
JD_DIS_U_X, JD_DIS_U_Y, JD_DIS_MASS, JD_DIS_GAMMA, JD_DIS_BIAS = range(JD_S, JD_S + 5)
JD_PUL_UA_X, JD_PUL_UA_Y, JD_PUL_UB_X, JD_PUL_UB_Y, JD_PUL_MASS = range(JD_S, JD_S + 5)
JD_MOU_MASS = JD_S
JD_MOU_GAMMA, JD_MOU_BETA, JD_MOU_C_X, JD_MOU_C_Y = range(JD_S + 4, JD_S + 8)
(JD_WHE_AX_X, JD_WHE_AX_Y, JD_WHE_AY_X, JD_WHE_AY_Y, JD_WHE_S_AX, JD_WHE_S_BX, JD_WHE_S_AY, JD_WHE_S_BY,
 JD_WHE_MASS, JD_WHE_MOTOR_MASS, JD_WHE_SPRING_MASS, JD_WHE_BIAS, JD_WHE_GAMMA) = range(JD_S, JD_S + 13)
JD_WEL_MASS = JD_S
JD_FRI_LMASS, JD_FRI_AMASS = JD_S, JD_S + 4
JD_ROP_U_X, JD_ROP_U_Y, JD_ROP_LENGTH, JD_ROP_MASS = range(JD_S, JD_S + 4)
(JD_GEA_JVAC_X, JD_GEA_JVAC_Y, JD_GEA_JVBD_X, JD_GEA_JVBD_Y, JD_GEA_JWA, JD_GEA_JWB, JD_GEA_JWC, JD_GEA_JWD,
 JD_GEA_MASS) = range(JD_RA_X, JD_RA_X + 9)
JD_GEA_V = JD_GEA_MASS + 1
JD_GEA_C = JD_GEA_V + 8
JD_GEA_A = JD_GEA_C + 8
JOINT_ND = 64


def pack_id(ia: int, ib: int, ta: int, tb: int) -> int:
    """Pack a contact-point id tuple (contact_util.py:78-79) into one int32."""
    u = (ia & 255) | ((ib & 255) << 8) | ((ta & 255) << 16) | ((tb & 255) << 24)
    return u - (1 << 32) if u >= (1 << 31) else u


def unpack_id(v: int):
    u = int(v) & 0xFFFFFFFF
    return (u & 255, (u >> 8) & 255, (u >> 16) & 255, (u >> 24) & 255)


class CBatch(ctypes.Structure):
    _fields_ = [
        ("n_worlds", ctypes.c_int32),
        ("body_off", ctypes.c_void_p), ("fix_off", ctypes.c_void_p),
        ("con_off", ctypes.c_void_p), ("move_off", ctypes.c_void_p),
        ("world_d", ctypes.c_void_p), ("world_i", ctypes.c_void_p),
        ("body_d", ctypes.c_void_p), ("body_i", ctypes.c_void_p),
        ("fix_d", ctypes.c_void_p), ("fix_i", ctypes.c_void_p),
        ("con_d", ctypes.c_void_p), ("con_i", ctypes.c_void_p),
        ("move", ctypes.c_void_p),
        ("joint_off", ctypes.c_void_p), ("joint_d", ctypes.c_void_p), ("joint_i", ctypes.c_void_p),
    ]


class CBatchOut(ctypes.Structure):
    _fields_ = [
        ("n_worlds", ctypes.c_int32), ("con_cap", ctypes.c_int32), ("move_cap", ctypes.c_int32),
        ("con_off", ctypes.c_void_p), ("move_off", ctypes.c_void_p),
        ("world_d", ctypes.c_void_p), ("world_i", ctypes.c_void_p),
        ("body_d", ctypes.c_void_p), ("body_i", ctypes.c_void_p),
        ("fix_d", ctypes.c_void_p), ("fix_i", ctypes.c_void_p),
        ("con_d", ctypes.c_void_p), ("con_i", ctypes.c_void_p),
        ("move", ctypes.c_void_p),
        ("joint_d", ctypes.c_void_p), ("joint_i", ctypes.c_void_p),
    ]


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data if a.size else 0


@dataclass
class BatchArrays:
    """All arrays of a batch (see include/b2g_layout.h)."""
    body_off: np.ndarray
    fix_off: np.ndarray
    con_off: np.ndarray
    move_off: np.ndarray
    world_d: np.ndarray
    world_i: np.ndarray
    body_d: np.ndarray
    body_i: np.ndarray
    fix_d: np.ndarray
    fix_i: np.ndarray
    con_d: np.ndarray
    con_i: np.ndarray
    move: np.ndarray
    joint_off: Optional[np.ndarray] = None  # n_worlds + 1 (None: no joints)
    joint_d: Optional[np.ndarray] = None
    joint_i: Optional[np.ndarray] = None
    _keep: list = field(default_factory=list, repr=False)

    def __post_init__(self):
        if self.joint_off is None:
            self.joint_off = np.zeros(int(np.asarray(self.world_d).reshape(-1, WORLD_ND).shape[0]) + 1, np.int32)
        if self.joint_d is None:
            self.joint_d = np.zeros((0, JOINT_ND))
        if self.joint_i is None:
            self.joint_i = np.zeros((0, JOINT_NI), np.int32)

    @property
    def n_joints(self) -> int:
        return int(np.asarray(self.joint_d).reshape(-1, JOINT_ND).shape[0])

    @property
    def n_worlds(self) -> int:
        return int(self.world_d.shape[0])

    @staticmethod
    def empty_world(gravity=(0.0, -10.0), flags=WF_ALLOW_SLEEP | WF_WARM_STARTING) -> "BatchArrays":
        wd = np.zeros((1, WORLD_ND), np.float64)
        wd[0, WD_GX], wd[0, WD_GY] = gravity
        wi = np.zeros((1, WORLD_NI), np.int32)
        wi[0, WI_FLAGS] = flags
        z = np.zeros(2, np.int32)
        return BatchArrays(z.copy(), z.copy(), z.copy(), z.copy(), wd, wi,
                           np.zeros((0, BODY_ND)), np.zeros((0, BODY_NI), np.int32),
                           np.zeros((0, FIX_ND)), np.zeros((0, FIX_NI), np.int32),
                           np.zeros((0, CON_ND)), np.zeros((0, CON_NI), np.int32),
                           np.zeros(0, np.int32))

    def normalized(self) -> "BatchArrays":
        """Contiguous arrays with the exact dtypes the C-ABI expects."""
        def f(a):
            return np.ascontiguousarray(a, dtype=np.float64)

        def i(a):
            return np.ascontiguousarray(a, dtype=np.int32)
        return BatchArrays(i(self.body_off), i(self.fix_off), i(self.con_off), i(self.move_off),
                           f(self.world_d).reshape(-1, WORLD_ND), i(self.world_i).reshape(-1, WORLD_NI),
                           f(self.body_d).reshape(-1, BODY_ND), i(self.body_i).reshape(-1, BODY_NI),
                           f(self.fix_d).reshape(-1, FIX_ND), i(self.fix_i).reshape(-1, FIX_NI),
                           f(self.con_d).reshape(-1, CON_ND), i(self.con_i).reshape(-1, CON_NI),
                           i(self.move), i(self.joint_off), f(self.joint_d).reshape(-1, JOINT_ND),
                           i(self.joint_i).reshape(-1, JOINT_NI))

    def as_struct(self) -> CBatch:
        b = self.normalized()
        s = CBatch(b.n_worlds, _ptr(b.body_off), _ptr(b.fix_off), _ptr(b.con_off), _ptr(b.move_off),
                   _ptr(b.world_d), _ptr(b.world_i), _ptr(b.body_d), _ptr(b.body_i),
                   _ptr(b.fix_d), _ptr(b.fix_i), _ptr(b.con_d), _ptr(b.con_i), _ptr(b.move),
                   _ptr(b.joint_off), _ptr(b.joint_d), _ptr(b.joint_i))
        self._keep = [b, s]
        return s

    def world(self, w: int) -> "BatchArrays":
        """A single-world batch holding world ``w`` (copies)."""
        b0, b1 = self.body_off[w], self.body_off[w + 1]
        f0, f1 = self.fix_off[w], self.fix_off[w + 1]
        c0, c1 = self.con_off[w], self.con_off[w + 1]
        m0, m1 = self.move_off[w], self.move_off[w + 1]
        j0, j1 = self.joint_off[w], self.joint_off[w + 1]
        return BatchArrays(np.array([0, b1 - b0], np.int32), np.array([0, f1 - f0], np.int32),
                           np.array([0, c1 - c0], np.int32), np.array([0, m1 - m0], np.int32),
                           self.world_d[w:w + 1].copy(), self.world_i[w:w + 1].copy(),
                           self.body_d[b0:b1].copy(), self.body_i[b0:b1].copy(),
                           self.fix_d[f0:f1].copy(), self.fix_i[f0:f1].copy(),
                           self.con_d[c0:c1].copy(), self.con_i[c0:c1].copy(),
                           self.move[m0:m1].copy(), np.array([0, j1 - j0], np.int32),
                           self.joint_d[j0:j1].copy(), self.joint_i[j0:j1].copy())

    def copy(self) -> "BatchArrays":
        return BatchArrays(*(getattr(self, n).copy() for n in ALL_FIELDS))


_FIELDS = ("body_off", "fix_off", "con_off", "move_off", "world_d", "world_i", "body_d",
           "body_i", "fix_d", "fix_i", "con_d", "con_i", "move")
_JOINT_FIELDS = ("joint_off", "joint_d", "joint_i")
ALL_FIELDS = _FIELDS + _JOINT_FIELDS


def load_batch(z, prefix: str = "") -> BatchArrays:
    """A batch from a mapping of arrays named prefix + field (an npz file); the joint arrays
    are optional."""
    names = getattr(z, "files", None) or list(z.keys())
    return BatchArrays(*[z[prefix + n] for n in _FIELDS],
                       *[z[prefix + n] if prefix + n in names else None for n in _JOINT_FIELDS])


def concat(worlds: List[BatchArrays]) -> BatchArrays:
    """Concatenate batches into one (world-local indices are kept)."""
    def offs(name):
        out = [0]
        for w in worlds:
            o = getattr(w, name)
            base = out[-1]
            out.extend(int(base + x) for x in o[1:])
        return np.array(out, np.int32)
    return BatchArrays(offs("body_off"), offs("fix_off"), offs("con_off"), offs("move_off"),
                       np.concatenate([w.world_d for w in worlds]).reshape(-1, WORLD_ND),
                       np.concatenate([w.world_i for w in worlds]).reshape(-1, WORLD_NI),
                       np.concatenate([w.body_d for w in worlds]).reshape(-1, BODY_ND),
                       np.concatenate([w.body_i for w in worlds]).reshape(-1, BODY_NI),
                       np.concatenate([w.fix_d for w in worlds]).reshape(-1, FIX_ND),
                       np.concatenate([w.fix_i for w in worlds]).reshape(-1, FIX_NI),
                       np.concatenate([w.con_d for w in worlds]).reshape(-1, CON_ND),
                       np.concatenate([w.con_i for w in worlds]).reshape(-1, CON_NI),
                       np.concatenate([w.move for w in worlds]).astype(np.int32), offs("joint_off"),
                       np.concatenate([w.joint_d for w in worlds]).reshape(-1, JOINT_ND),
                       np.concatenate([w.joint_i for w in worlds]).reshape(-1, JOINT_NI).astype(np.int32))


class OutBuffers:
    """Writable output arrays for a batch (``b2g_batch_out``)."""

    def __init__(self, like: BatchArrays, con_cap: Optional[int] = None, move_cap: Optional[int] = None):
        nw = like.n_worlds
        nb = like.body_d.shape[0]
        nf = like.fix_d.shape[0]
        self.con_cap = int(con_cap if con_cap is not None else max(16, 2 * like.con_d.shape[0] + 8 * nf))
        self.move_cap = int(move_cap if move_cap is not None else nf + 16)
        self.con_off = np.zeros(nw + 1, np.int32)
        self.move_off = np.zeros(nw + 1, np.int32)
        self.world_d = like.world_d.copy()
        self.world_i = like.world_i.copy()
        self.body_d = np.zeros((nb, BODY_ND))
        self.body_i = np.zeros((nb, BODY_NI), np.int32)
        self.fix_d = like.fix_d.copy()
        self.fix_i = like.fix_i.copy()
        self.con_d = np.zeros((self.con_cap, CON_ND))
        self.con_i = np.zeros((self.con_cap, CON_NI), np.int32)
        self.move = np.zeros(self.move_cap, np.int32)
        self.body_off = like.body_off.copy()
        self.fix_off = like.fix_off.copy()
        self.joint_off = np.ascontiguousarray(like.joint_off, dtype=np.int32).copy()
        self.joint_d = np.ascontiguousarray(like.joint_d, dtype=np.float64).reshape(-1, JOINT_ND).copy()
        self.joint_i = np.ascontiguousarray(like.joint_i, dtype=np.int32).reshape(-1, JOINT_NI).copy()
        self.struct = CBatchOut(nw, self.con_cap, self.move_cap, _ptr(self.con_off), _ptr(self.move_off),
                                _ptr(self.world_d), _ptr(self.world_i), _ptr(self.body_d), _ptr(self.body_i),
                                _ptr(self.fix_d), _ptr(self.fix_i), _ptr(self.con_d), _ptr(self.con_i),
                                _ptr(self.move), _ptr(self.joint_d), _ptr(self.joint_i))

    def to_batch(self) -> BatchArrays:
        nc = int(self.con_off[-1])
        nm = int(self.move_off[-1])
        return BatchArrays(self.body_off.copy(), self.fix_off.copy(), self.con_off.copy(),
                           self.move_off.copy(), self.world_d.copy(), self.world_i.copy(),
                           self.body_d.copy(), self.body_i.copy(), self.fix_d.copy(), self.fix_i.copy(),
                           self.con_d[:nc].copy(), self.con_i[:nc].copy(), self.move[:nm].copy(),
                           self.joint_off.copy(), self.joint_d.copy(), self.joint_i.copy())
