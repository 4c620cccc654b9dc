"""The oracle's queries and ray casts against the reference's recorded results and closed forms.

Fixtures (tests/golden/make_golden.py, recorded from pypybox2d's pure-Python
flavour):
  * shape_raycast.npz: 360 Shape.ray_cast calls (circles, polygons, edges at
    random transforms; rays through the shape, past it, and from inside);
  * queries_mixed24.npz / queries_filtered.npz: a reference world after 120 /
    150 steps, with World.query_aabb over 200 random boxes (the fixtures
    yielded) and World.ray_cast over 200 random segments with every hit
    continued unclipped (fixture, fraction, point, normal).
The oracle must reproduce every one bit for bit.  The closed forms pin the
meaning: a ray along the x axis meets a unit circle at a known fraction, a
ray starting inside a polygon reports nothing, and an edge's hit normal is
returned in the edge's own frame (shapes.py:655-702 never rotates it back).
"""
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN
from oracle import oracle as O
from pypybox2d_amd import builder, layout as L


def test_shape_ray_casts_match_the_reference_bit_exactly():
    z = np.load(os.path.join(GOLDEN, "shape_raycast.npz"))
    n = z["fd"].shape[0]
    hits = 0
    for k in range(n):
        hit, f, nx, ny = O.shape_ray_cast(z["fd"][k], z["fi"][k], z["xf"][k], z["seg"][k])
        assert hit == bool(z["hit"][k]), k
        if hit:
            hits += 1
            assert (f, nx, ny) == tuple(z["out"][k]), k
    kinds = z["fi"][:, L.FI_SHAPE]
    for t in (L.SHAPE_CIRCLE, L.SHAPE_POLYGON, L.SHAPE_EDGE):  # hits and misses of every kind
        assert 0 < int(z["hit"][kinds == t].sum()) < int((kinds == t).sum())
    assert hits > n // 3


@pytest.mark.parametrize("name", ["queries_mixed24.npz", "queries_filtered.npz"])
def test_world_queries_and_ray_casts_match_the_reference(name):
    z = np.load(os.path.join(GOLDEN, name))
    o = O.Oracle(L.BatchArrays(*[z["init_" + n] for n in L._FIELDS]))
    nonempty = 0
    for q in range(z["boxes"].shape[0]):
        got = o.query_aabb(0, z["boxes"][q])
        want = z["q_fix"][z["q_off"][q]:z["q_off"][q + 1]]
        assert np.array_equal(got, want), q
        nonempty += int(len(want) > 0)
    hit_rays = 0
    for r in range(z["segs"].shape[0]):
        f, d = o.ray_cast(0, z["segs"][r])
        a, b = z["r_off"][r], z["r_off"][r + 1]
        assert np.array_equal(f, z["r_fix"][a:b]), r
        assert np.array_equal(d, z["r_dat"][a:b]), r
        hit_rays += int(b > a)
    assert nonempty > 50 and hit_rays > 50


def one_shape_world(shape, position, angle=0.0):
    sc = {"gravity": (0.0, -10.0), "do_sleep": True, "warm_starting": True,
          "bodies": [{"type": "static", "position": position, "angle": angle, "linear_velocity": (0.0, 0.0),
                      "angular_velocity": 0.0, "fixtures": [{"shape": shape, "density": 0.0, "friction": 0.2,
                                                             "restitution": 0.0}]}]}
    return O.Oracle(builder.scene_batch(sc))


def test_closed_forms():
    # a ray along the x axis meets the circle of radius 1 at (5, 0) at x = 4
    o = one_shape_world(("circle", 1.0), (5.0, 0.0))
    f, d = o.ray_cast(0, (0.0, 0.0, 10.0, 0.0))
    assert list(f) == [0] and d[0, 0] == 0.4 and tuple(d[0, 1:]) == (4.0, 0.0, -1.0, 0.0)
    # the same ray clipped before the circle, or missing it above, reports nothing
    assert o.ray_cast(0, (0.0, 0.0, 3.9, 0.0))[0].size == 0
    assert o.ray_cast(0, (0.0, 1.5, 10.0, 1.5))[0].size == 0
    # a box: the entering face's normal; from inside: nothing
    o = one_shape_world(("box", 1.0, 0.5), (0.0, 0.0))
    f, d = o.ray_cast(0, (0.0, 5.0, 0.0, -5.0))
    assert list(f) == [0] and abs(d[0, 0] - 0.45) < 1e-15 and tuple(d[0, 3:]) == (0.0, 1.0)
    assert o.ray_cast(0, (0.1, 0.1, 0.0, 5.0))[0].size == 0
    # an edge rotated by 90 degrees: the world ray hits it along x, the normal is the local one
    o = one_shape_world(("edge", (-1.0, 0.0), (1.0, 0.0)), (0.0, 0.0), angle=math.pi / 2)
    f, d = o.ray_cast(0, (-3.0, 0.2, 3.0, 0.2))
    assert list(f) == [0] and abs(d[0, 0] - 0.5) < 1e-15
    assert np.allclose(d[0, 3:], (0.0, 1.0)) or np.allclose(d[0, 3:], (0.0, -1.0))  # local y, not world x
    # query: the fat AABB (shape AABB grown by 0.1) is what a box query sees, inclusively
    o = one_shape_world(("box", 1.0, 1.0), (0.0, 0.0))
    edge = 1.0 + 0.01 + 0.1  # half-extent + polygon radius + AABB extension
    assert list(o.query_aabb(0, (edge, -0.5, edge + 1.0, 0.5))) == [0]
    assert list(o.query_aabb(0, (edge + 1e-9, -0.5, edge + 1.0, 0.5))) == []
    with pytest.raises(ValueError):
        o.ray_cast(0, (1.0, 1.0, 1.0, 1.0))
