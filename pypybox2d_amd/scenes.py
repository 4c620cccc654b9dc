"""Synthetic scenes of the benchmark and parity workloads, as plain data.

A scene is a dict ``{"gravity", "do_sleep", "warm_starting", "bodies"}``;
every body is a dict with ``type`` ("static" / "dynamic" / "kinematic"), ``position``,
``angle``, ``linear_velocity``, ``angular_velocity`` and a list of
``fixtures`` (``shape``, ``density``, ``friction``, ``restitution`` and, when
not the defaults, ``category``, ``mask`` and ``group``); a body may carry
``bullet``, and the scene ``continuous`` (continuous physics on).  Shapes
are tuples: ``("box", hx, hy)``, ``("circle", r)``, ``("polygon", [(x, y),
...])`` (convex, counter-clockwise) or ``("edge", (x1, y1), (x2, y2))``.

The same spec builds a pypybox2d world (through its public API) or a world of
this package, so both start from the same inputs.

Workloads (gravity (0, -10), dt = 1/60, 8 velocity / 3 position iterations):
  pyramid      10-row box pyramid on an edge ground (55 boxes)
  stacks       5 columns of 10 boxes on a static box ground, seeded x jitter
  mixed_world  one world of boxes / circles / convex polygons above an edge
  mixed_batch  many mixed worlds with body counts uniform in [32, 512]
  pyramids     N pyramids side by side on their own ground edges
  lognormal_batch  many mixed worlds with lognormal body counts
  filtered_world   mixed bodies with category/mask/group filtering and a kinematic platform
  projectiles  fast bodies (some bullets) shot at a thin wall and into the ground
               (continuous physics)
  joint_bridge    planks pinned end to end by revolute joints, shapes dropped onto them
  joint_machines  one mechanism per joint type side by side (limits, motors, springs, a
                  pulley, a rope, a car on wheel joints, a friction joint, a mouse joint)
  joint_gears     gear joints over revolute and prismatic joints

A scene may carry ``joints``: each a dict with ``type`` (revolute, prismatic, distance,
pulley, mouse, gear, wheel, weld, friction, rope), ``bodies`` (indices into the body list;
one body for a mouse joint) or, for a gear joint, ``joints`` (indices into the joint list),
and the keyword arguments of the reference's create_<type>_joint.
"""
from __future__ import annotations

import math
from typing import List

import numpy as np


def _body(type_, position, fixtures, angle=0.0, linear_velocity=(0.0, 0.0), angular_velocity=0.0, bullet=False):
    b = {"type": type_, "position": tuple(position), "angle": angle,
         "linear_velocity": tuple(linear_velocity), "angular_velocity": angular_velocity,
         "fixtures": fixtures}
    if bullet:
        b["bullet"] = True
    return b


def _fix(shape, density=0.0, friction=0.2, restitution=0.0, category=0x0001, mask=0xFFFF, group=0):
    f = {"shape": shape, "density": density, "friction": friction, "restitution": restitution}
    if (category, mask, group) != (0x0001, 0xFFFF, 0):
        f.update(category=category, mask=mask, group=group)
    return f


def fixture_filter(f):
    """(category_bits, mask_bits, group_index) of a fixture spec (defaults as Fixture's)."""
    return f.get("category", 0x0001), f.get("mask", 0xFFFF), f.get("group", 0)


def pyramid(rows: int = 10, x0: float = 0.0, ground_half: float = 40.0, with_ground: bool = True):
    """Box pyramid: row i has rows-i boxes at x = -7 + 0.5625 i + 1.125 j, y = 0.75 + i."""
    bodies = []
    if with_ground:
        bodies.append(_body("static", (0.0, 0.0),
                            [_fix(("edge", (x0 - ground_half, 0.0), (x0 + ground_half, 0.0)))]))
    for i in range(rows):
        for j in range(rows - i):
            x = -7 + 0.5625 * i + 1.125 * j
            if x0 != 0.0:
                x = x + x0
            bodies.append(_body("dynamic", (x, 0.75 + 1.0 * i),
                                [_fix(("box", 0.5, 0.5), density=5.0, friction=0.6)]))
    return {"gravity": (0.0, -10.0), "do_sleep": True, "warm_starting": True, "bodies": bodies}


def stacks(seed: int = 0, columns=(-6.0, -3.0, 0.0, 3.0, 6.0), rows: int = 10):
    """Vertical stacks of half-extent 0.5 boxes on a static 40x1 box ground at (0, -1)."""
    rng = np.random.default_rng(seed)
    bodies = [_body("static", (0.0, -1.0), [_fix(("box", 40.0, 1.0))])]
    for cx in columns:
        for k in range(rows):
            x = float(cx) + float(rng.uniform(-0.01, 0.01))
            bodies.append(_body("dynamic", (x, 0.75 + 1.1 * k),
                                [_fix(("box", 0.5, 0.5), density=1.0, friction=0.3)]))
    return {"gravity": (0.0, -10.0), "do_sleep": True, "warm_starting": True, "bodies": bodies}


def _random_shape(rng):
    kind = int(rng.choice(3, p=[0.4, 0.3, 0.3]))
    if kind == 0:
        hx, hy = (float(v) for v in rng.uniform(0.2, 1.0, 2))
        return ("box", hx, hy), math.hypot(hx, hy)
    if kind == 1:
        r = float(rng.uniform(0.2, 1.0))
        return ("circle", r), r
    k = int(rng.integers(3, 9))
    rad = float(rng.uniform(0.3, 1.0))
    ang = 2.0 * np.pi * np.arange(k) / k + rng.uniform(-0.3, 0.3, k) * (np.pi / k)
    ang = np.sort(ang)
    verts = [(float(rad * math.cos(a)), float(rad * math.sin(a))) for a in ang]
    return ("polygon", verts), rad


def mixed_world(rng, n_bodies: int, half_width: float = 20.0, height: float = None):
    """Boxes / circles / convex polygons (40/30/30) placed without overlap above an edge.

    Centres are rejection-sampled in [-half_width, half_width] x [1, 1 + height]
    with non-overlapping bounding circles (0.05 m apart), largest bodies
    first.  The default region is 40 x 40 and grows in height for large worlds
    so that the bounding circles cover at most ~40% of it.  Each body then gets
    a uniform random angle, v ~ N(0, 1) per axis and w ~ N(0, 0.5).
    """
    if height is None:
        height = max(40.0, 4.7 * n_bodies / (2.0 * half_width))
    shapes = []
    for _ in range(n_bodies):
        shape, R = _random_shape(rng)
        density = float(rng.uniform(0.5, 2.0))
        friction = float(rng.uniform(0.2, 0.8))
        shapes.append((shape, R, density, friction))
    cell = 3.0
    grid = {}
    pos = [None] * n_bodies
    for i in sorted(range(n_bodies), key=lambda k: -shapes[k][1]):
        R = shapes[i][1]
        for _try in range(100000):
            x = float(rng.uniform(-half_width, half_width))
            y = float(rng.uniform(1.0, 1.0 + height))
            if y - R < 0.05:
                continue
            gx, gy = int(math.floor(x / cell)), int(math.floor(y / cell))
            ok = True
            for dx in (-1, 0, 1):
                for dy in (-1, 0, 1):
                    for (px, py, pr) in grid.get((gx + dx, gy + dy), ()):
                        if (px - x) ** 2 + (py - y) ** 2 < (pr + R + 0.05) ** 2:
                            ok = False
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if ok:
                break
        else:
            raise RuntimeError("could not place body")
        grid.setdefault((gx, gy), []).append((x, y, R))
        pos[i] = (x, y)
    bodies = [_body("static", (0.0, 0.0), [_fix(("edge", (-40.0, 0.0), (40.0, 0.0)))])]
    for i, (shape, R, density, friction) in enumerate(shapes):
        v = rng.normal(0.0, 1.0, 2)
        w = float(rng.normal(0.0, 0.5))
        angle = float(rng.uniform(-np.pi, np.pi))
        bodies.append(_body("dynamic", pos[i], [_fix(shape, density=density, friction=friction)],
                            angle=angle, linear_velocity=(float(v[0]), float(v[1])), angular_velocity=w))
    return {"gravity": (0.0, -10.0), "do_sleep": True, "warm_starting": True, "bodies": bodies}


def mixed_batch_counts(n_worlds: int = 4096, seed: int = 1, lo: int = 32, hi: int = 512) -> np.ndarray:
    """Body counts of the uniform batch: uniform integers in [lo, hi]."""
    return np.random.default_rng(seed).integers(lo, hi + 1, n_worlds)


def lognormal_batch_counts(n_worlds: int = 16384, seed: int = 2, median: float = 64.0, sigma: float = 0.75,
                           lo: int = 16, hi: int = 1024) -> np.ndarray:
    """Body counts clip(round(median * exp(sigma N(0,1))), lo, hi)."""
    rng = np.random.default_rng(seed)
    return np.clip(np.round(median * np.exp(sigma * rng.normal(size=n_worlds))), lo, hi).astype(int)


def batch_world(seed: int, w: int, n_bodies: int) -> dict:
    """World w of a batch: its own random stream, seeded by (seed, w)."""
    return mixed_world(np.random.default_rng([seed, w]), int(n_bodies))


def mixed_batch(n_worlds: int = 4096, seed: int = 1, lo: int = 32, hi: int = 512) -> List[dict]:
    """Independent mixed worlds with body counts uniform in [lo, hi]."""
    counts = mixed_batch_counts(n_worlds, seed, lo, hi)
    return [batch_world(seed, w, n) for w, n in enumerate(counts)]


def lognormal_batch(n_worlds: int = 16384, seed: int = 2, median: float = 64.0, sigma: float = 0.75,
                    lo: int = 16, hi: int = 1024) -> List[dict]:
    """Mixed worlds with lognormal body counts (see lognormal_batch_counts)."""
    counts = lognormal_batch_counts(n_worlds, seed, median, sigma, lo, hi)
    return [batch_world(seed, w, n) for w, n in enumerate(counts)]


def filtered_world(seed: int = 5, n_bodies: int = 36):
    """Mixed bodies with collision filtering and a kinematic platform.

    Fixtures take category bits 1, 2, 4 or 8 in turn.  Category 2 ignores category 4 and
    4 ignores 2 (mask bits); category 8 ignores its own category, but its bodies are in group
    +1 and collide with each other anyway (a positive group overrides the masks), and every
    seventh body (from the fourth) is in group -1, whose members never collide with each
    other.  A kinematic box
    platform (6 x 0.3 half-extents) at (0, 5) moves right at 1 m/s and turns at 0.2 rad/s;
    bodies fall on it and on the ground edge."""
    rng = np.random.default_rng(seed)
    bodies = [_body("static", (0.0, 0.0), [_fix(("edge", (-30.0, 0.0), (30.0, 0.0)))]),
              _body("kinematic", (0.0, 5.0), [_fix(("box", 6.0, 0.3), friction=0.5)],
                    linear_velocity=(1.0, 0.0), angular_velocity=0.2)]
    cats = (0x0001, 0x0002, 0x0004, 0x0008)
    masks = {0x0001: 0xFFFF, 0x0002: 0xFFFF & ~0x0004, 0x0004: 0xFFFF & ~0x0002, 0x0008: 0xFFFF & ~0x0008}
    for i in range(n_bodies):
        shape, _ = _random_shape(rng)
        if shape[0] == "box":
            shape = ("box", min(shape[1], 0.6), min(shape[2], 0.6))
        elif shape[0] == "circle":
            shape = ("circle", min(shape[1], 0.6))
        cat = cats[i % 4]
        group = -1 if i % 7 == 3 else (1 if cat == 0x0008 else 0)
        x = float(rng.uniform(-3.0, 3.0)) if cat == 0x0008 else float(rng.uniform(-6.0, 6.0))
        y = 7.0 + 1.4 * i
        bodies.append(_body("dynamic", (x, y), [_fix(shape, density=float(rng.uniform(0.5, 2.0)),
                                                     friction=float(rng.uniform(0.2, 0.8)), category=cat,
                                                     mask=masks[cat], group=group)],
                            angle=float(rng.uniform(-np.pi, np.pi)),
                            linear_velocity=(float(rng.normal(0, 0.5)), 0.0)))
    return {"gravity": (0.0, -10.0), "do_sleep": True, "warm_starting": True, "bodies": bodies}


def projectiles(seed: int = 3):
    """Fast bodies that tunnel without continuous physics: a thin static wall (0.1 m half
    width) at x = 10 on an edge ground, a column of dynamic boxes in front of it, and 10
    projectiles (circles, boxes, a triangle; every other one a bullet) shot at it at 60-150 m/s
    (1-2.5 m per step), plus 3 bodies thrown into the ground at 50-90 m/s and a kinematic
    paddle sweeping through."""
    rng = np.random.default_rng(seed)
    bodies = [_body("static", (0.0, 0.0), [_fix(("edge", (-30.0, 0.0), (30.0, 0.0)))]),
              _body("static", (10.0, 4.0), [_fix(("box", 0.1, 4.0))]),
              _body("kinematic", (0.0, 9.0), [_fix(("box", 2.0, 0.15))], linear_velocity=(0.0, -1.5),
                    angular_velocity=1.0)]
    for k in range(4):  # dynamic boxes against the wall: bullets hit them (bullet vs dynamic)
        bodies.append(_body("dynamic", (9.3, 0.5 + 1.01 * k), [_fix(("box", 0.5, 0.5), density=1.0, friction=0.6)]))
    shapes = [("circle", 0.2), ("box", 0.2, 0.15), ("polygon", [(-0.2, -0.15), (0.25, -0.1), (0.0, 0.2)]),
              ("circle", 0.12), ("box", 0.1, 0.3)]
    for k in range(10):
        sh = shapes[k % len(shapes)]
        speed = float(rng.uniform(60.0, 150.0))
        y = 0.6 + 0.75 * k
        ang = float(rng.uniform(-0.05, 0.05))
        bodies.append(_body("dynamic", (-4.0 - 1.5 * k, y), [_fix(sh, density=float(rng.uniform(0.5, 3.0)),
                                                                     friction=0.3, restitution=0.1 * (k % 3))],
                            angle=float(rng.uniform(-1, 1)), linear_velocity=(speed * math.cos(ang), speed * math.sin(ang)),
                            angular_velocity=float(rng.uniform(-5, 5)), bullet=(k % 2 == 0)))
    for k in range(3):
        bodies.append(_body("dynamic", (-12.0 + 4.0 * k, 6.0), [_fix(shapes[k], density=1.0, friction=0.4)],
                            linear_velocity=(float(rng.uniform(-5, 5)), -float(rng.uniform(50, 90)))))
    return {"gravity": (0.0, -10.0), "do_sleep": True, "warm_starting": True, "bodies": bodies}


def _joint(kind, *bodies, **params):
    j = {"type": kind, "bodies": tuple(bodies)}
    j.update(params)
    return j


def joint_bridge(planks: int = 30):
    """The reference testbed's bridge (testbed/bridge.py): planks of half extents 0.5 x 0.125
    (density 20, friction 0.2) pinned end to end by revolute joints between the ground and the
    ground again, with triangles and circles dropped onto it."""
    bodies = [_body("static", (0.0, 0.0), [_fix(("edge", (-40.0, 0.0), (40.0, 0.0)))])]
    joints = []
    prev = 0
    for i in range(planks):
        bodies.append(_body("dynamic", (-14.5 + i, 5.0), [_fix(("box", 0.5, 0.125), density=20.0, friction=0.2)]))
        joints.append(_joint("revolute", prev, len(bodies) - 1, anchor=(-15.0 + i, 5.0)))
        prev = len(bodies) - 1
    joints.append(_joint("revolute", prev, 0, anchor=(-15.0 + planks, 5.0)))
    for i in range(2):
        bodies.append(_body("dynamic", (-8.0 + 8.0 * i, 12.0),
                            [_fix(("polygon", [(-0.5, 0.0), (0.5, 0.0), (0.0, 1.5)]), density=1.0)]))
    for i in range(3):
        bodies.append(_body("dynamic", (-6.0 + 6.0 * i, 10.0), [_fix(("circle", 0.5), density=1.0)]))
    return {"gravity": (0.0, -10.0), "do_sleep": True, "warm_starting": True, "bodies": bodies, "joints": joints}


def joint_machines():
    """Every joint type in one world, each mechanism on its own stretch of a long ground edge:
    a pendulum between revolute limits, a motor-driven wheel, a revolute joint with equal limits,
    a prismatic elevator with limits and a motor, a slanted prismatic slide, a welded cantilever,
    a chain held by a rope joint, rigid and soft distance joints, a pulley, a car on wheel joints
    (one motor), a box held back by a friction joint and a box dragged by a mouse joint."""
    pi = math.pi
    bodies = [_body("static", (0.0, 0.0), [_fix(("edge", (-80.0, 0.0), (80.0, 0.0)), friction=0.6)])]
    joints = []

    def add(b):
        bodies.append(b)
        return len(bodies) - 1
    # pendulum between limits (starts swinging)
    p = add(_body("dynamic", (-60.0, 6.0), [_fix(("box", 0.25, 1.5), density=2.0)], angular_velocity=8.0))
    joints.append(_joint("revolute", 0, p, anchor=(-60.0, 7.5), lower_angle=-0.15 * pi, upper_angle=0.15 * pi,
                         limit_enabled=True))
    p = add(_body("dynamic", (-56.0, 6.0), [_fix(("box", 0.25, 1.5), density=2.0)], angular_velocity=-8.0))
    joints.append(_joint("revolute", 0, p, anchor=(-56.0, 7.5), lower_angle=-0.15 * pi, upper_angle=0.15 * pi,
                         limit_enabled=True))
    # motor-driven wheel with a box on top
    wh = add(_body("dynamic", (-52.0, 3.0), [_fix(("circle", 1.0), density=1.0, friction=0.8)]))
    joints.append(_joint("revolute", 0, wh, anchor=(-52.0, 3.0), motor_enabled=True, motor_speed=4.0,
                         max_motor_torque=50.0))
    add(_body("dynamic", (-52.0, 4.6), [_fix(("box", 0.5, 0.5), density=1.0, friction=0.8)]))
    # equal limits: held at 0.3 rad
    e = add(_body("dynamic", (-46.0, 5.0), [_fix(("box", 1.0, 0.2), density=1.0)]))
    joints.append(_joint("revolute", 0, e, anchor=(-47.0, 5.0), lower_angle=0.3, upper_angle=0.3,
                         limit_enabled=True))
    # prismatic elevator (limits 0..4, motor up) carrying a box
    el = add(_body("dynamic", (-38.0, 1.0), [_fix(("box", 2.0, 0.2), density=2.0, friction=0.6)]))
    joints.append(_joint("prismatic", 0, el, anchor=(-38.0, 1.0), axis=(0.0, 1.0), lower_limit=0.0,
                         upper_limit=4.0, limit_enabled=True, motor_enabled=True, motor_speed=1.5,
                         max_motor_force=2000.0))
    add(_body("dynamic", (-38.0, 1.8), [_fix(("box", 0.5, 0.5), density=1.0, friction=0.6)]))
    # slanted slide, no limit or motor
    sl = add(_body("dynamic", (-30.0, 4.0), [_fix(("box", 0.5, 0.5), density=1.0)]))
    joints.append(_joint("prismatic", 0, sl, anchor=(-30.0, 4.0), axis=(1.0, -0.4)))
    # welded cantilever of six planks
    prev = 0
    for i in range(6):
        b = add(_body("dynamic", (-21.5 + i, 5.0), [_fix(("box", 0.5, 0.125), density=20.0, friction=0.2)]))
        joints.append(_joint("weld", prev, b, anchor=(-22.0 + i, 5.0)))
        prev = b
    # chain of planks under a rope joint (testbed/ropes.py)
    prev = 0
    for i in range(8):
        b = add(_body("dynamic", (-9.5 + i, 15.0), [_fix(("box", 0.5, 0.125), density=20.0, friction=0.2)]))
        joints.append(_joint("revolute", prev, b, anchor=(-10.0 + i, 15.0)))
        prev = b
    joints.append(_joint("rope", 0, prev, local_anchor_a=(-10.0, 15.0), local_anchor_b=(0.5, 0.0),
                         max_length=7.01))
    # distance joints: a rigid pair of circles hanging from the ground, and soft springs
    c1 = add(_body("dynamic", (4.0, 10.0), [_fix(("circle", 0.4), density=1.0)]))
    c2 = add(_body("dynamic", (5.5, 9.0), [_fix(("circle", 0.4), density=1.0)]))
    joints.append(_joint("distance", 0, c1, anchor_a=(4.0, 14.0), anchor_b=(4.0, 10.0)))
    joints.append(_joint("distance", c1, c2, anchor_a=(4.0, 10.0), anchor_b=(5.5, 9.0), collide_connected=True))
    s1 = add(_body("dynamic", (10.0, 9.0), [_fix(("box", 0.4, 0.4), density=1.0)]))
    s2 = add(_body("dynamic", (12.0, 9.0), [_fix(("box", 0.4, 0.4), density=1.0)]))
    joints.append(_joint("distance", 0, s1, anchor_a=(10.0, 13.0), anchor_b=(10.0, 9.4), frequency=3.0,
                         damping_ratio=0.4))
    joints.append(_joint("distance", s1, s2, anchor_a=(10.0, 9.0), anchor_b=(12.0, 9.0), frequency=4.0,
                         damping_ratio=0.1))
    # pulley (testbed/pulley.py, smaller)
    pa = add(_body("dynamic", (20.0, 8.0), [_fix(("box", 0.5, 1.0), density=5.0)]))
    pb = add(_body("dynamic", (26.0, 8.0), [_fix(("box", 0.5, 1.0), density=5.0)]))
    joints.append(_joint("pulley", pa, pb, ground_anchor_a=(20.0, 14.0), ground_anchor_b=(26.0, 14.0),
                         anchor_a=(20.0, 9.0), anchor_b=(26.0, 9.0), ratio=1.5))
    # car: chassis on two wheel joints, rear wheel driven
    ch = add(_body("dynamic", (40.0, 1.4), [_fix(("box", 1.5, 0.4), density=1.0)]))
    for k, x in enumerate((38.8, 41.2)):
        w_ = add(_body("dynamic", (x, 0.6), [_fix(("circle", 0.4), density=1.0, friction=0.9)]))
        joints.append(_joint("wheel", ch, w_, anchor=(x, 0.6), axis=(0.0, 1.0), frequency=4.0, damping_ratio=0.7,
                             motor_enabled=(k == 0), motor_speed=-8.0, max_motor_torque=20.0))
    # friction joint holding back a sliding box
    fr = add(_body("dynamic", (55.0, 0.5), [_fix(("box", 0.5, 0.5), density=1.0, friction=0.1)],
                   linear_velocity=(6.0, 0.0), angular_velocity=1.0))
    joints.append(_joint("friction", 0, fr, local_anchor_a=(55.0, 0.0), local_anchor_b=(0.0, 0.0), max_force=8.0,
                         max_torque=2.0))
    # mouse joint dragging a box toward a point
    mo = add(_body("dynamic", (65.0, 3.0), [_fix(("box", 0.5, 0.5), density=1.0)]))
    joints.append({"type": "mouse", "bodies": (mo,), "target": (68.0, 7.0), "max_force": 300.0, "frequency": 5.0,
                   "damping_ratio": 0.7})
    return {"gravity": (0.0, -10.0), "do_sleep": True, "warm_starting": True, "bodies": bodies, "joints": joints}


def joint_gears():
    """The reference testbed's gears (testbed/gear.py): a box pinned to a static circle and to a
    dynamic circle, coupled by a gear joint; two circles on revolute joints and a box on a
    prismatic joint to the ground, coupled by gear joints (revolute-revolute and
    revolute-prismatic)."""
    bodies = [_body("static", (0.0, 0.0), [_fix(("edge", (50.0, 0.0), (-50.0, 0.0)))])]
    bodies.append(_body("static", (10.0, 9.0), [_fix(("circle", 1.0))]))                    # 1
    bodies.append(_body("dynamic", (10.0, 8.0), [_fix(("box", 0.5, 0.5), density=5.0)]))    # 2
    bodies.append(_body("dynamic", (10.0, 6.0), [_fix(("circle", 2.0), density=5.0)]))     # 3
    bodies.append(_body("dynamic", (-3.0, 12.0), [_fix(("circle", 1.0), density=5.0)]))    # 4
    bodies.append(_body("dynamic", (0.0, 12.0), [_fix(("circle", 2.0), density=5.0)]))     # 5
    bodies.append(_body("dynamic", (2.5, 12.0), [_fix(("box", 0.5, 0.5), density=5.0)]))   # 6
    joints = [_joint("revolute", 2, 1, anchor=(10.0, 9.0)),                                 # 0
              _joint("revolute", 2, 3, anchor=(10.0, 6.0)),                                 # 1
              {"type": "gear", "joints": (0, 1), "ratio": 2.0},                             # 2
              _joint("revolute", 0, 4, anchor=(-3.0, 12.0)),                                # 3
              _joint("revolute", 0, 5, anchor=(0.0, 12.0)),                                 # 4
              _joint("prismatic", 0, 6, anchor=(2.5, 12.0), axis=(0.0, 1.0), lower_limit=-5.0,
                     upper_limit=5.0, limit_enabled=True),                                   # 5
              {"type": "gear", "joints": (3, 4), "ratio": 2.0},                             # 6
              {"type": "gear", "joints": (4, 5), "ratio": -1.0 / 2.0}]                      # 7
    return {"gravity": (0.0, -10.0), "do_sleep": True, "warm_starting": True, "bodies": bodies, "joints": joints}


def listener_filter_verdict(index_a: int, bits_a, index_b: int, bits_b) -> bool:
    """The custom contact filter of the listener scenes, as a function of the two fixtures'
    indices and (category, mask, group) bits: World.default_contact_filter on the bits, and
    then a rule on the indices that is not symmetric ((a + 2 b) mod 5 != 0), so that a call
    with the fixtures swapped would often give another answer."""
    ca, ma, ga = bits_a
    cb, mb, gb = bits_b
    if ga == gb and ga != 0:
        ok = ga > 0
    else:
        ok = (ma & cb) != 0 and (ca & mb) != 0
    return ok and (index_a + 2 * index_b) % 5 != 0


def listener_disable_rule(step: int, index_a: int, index_b: int) -> bool:
    """The pre_solve of the listener scenes sets contact.enabled = False for the step when this
    holds for the step number and the contact's fixtures' indices."""
    return (step + index_a + 2 * index_b) % 6 == 0


def pyramids(count: int = 200, spacing: float = 30.0):
    """One world holding `count` 10-row pyramids, each on its own static ground edge."""
    bodies = []
    for k in range(count):
        x0 = spacing * k
        bodies.append(_body("static", (0.0, 0.0), [_fix(("edge", (-12.0 + x0, 0.0), (12.0 + x0, 0.0)))]))
    for k in range(count):
        bodies.extend(pyramid(x0=spacing * k, with_ground=False)["bodies"])
    return {"gravity": (0.0, -10.0), "do_sleep": True, "warm_starting": True, "bodies": bodies}


def dynamic_body_count(scene) -> int:
    return sum(1 for b in scene["bodies"] if b["type"] == "dynamic")
