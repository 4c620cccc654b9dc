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
