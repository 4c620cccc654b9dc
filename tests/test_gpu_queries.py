"""GPU: World.query_aabb and World.ray_cast on the device's LBVH, and the large worlds' pair
search by tree traversal, through the C-ABI.

Bars: query results and ray-cast hits (fixture, fraction, point, normal) equal the oracle's
brute force bit for bit on the same state, and the reference's recorded results on its
recorded worlds (tests/golden/queries_*.npz); the closest hit is the all-hits list's first
entry by (fraction, fixture); the tree pair search gives exactly the contacts of the
all-pairs search (bit-identical states over many steps).
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN
from oracle import oracle as O
from pypybox2d_amd import _native as N
from pypybox2d_amd import builder, layout as L, scenes

pytestmark = pytest.mark.gpu

DT = 1.0 / 60


def engine_with(state, prec="fp64", **params):
    e = N.Engine(prec)
    for k, v in params.items():
        e.set_param(k, v)
    e.upload(state)
    return e


def split_rows(rows, data, n):
    """Per ray: (fixtures, data rows) from rows sorted by ray."""
    out = []
    for r in range(n):
        m = rows[:, 0] == r
        out.append((rows[m, 1], data[m]))
    return out


@pytest.mark.parametrize("name", ["queries_mixed24.npz", "queries_filtered.npz"])
def test_queries_and_ray_casts_match_the_recorded_reference(name):
    z = np.load(os.path.join(GOLDEN, name))
    st = L.BatchArrays(*[z["init_" + n] for n in L._FIELDS])
    e = engine_with(st)
    nq, nr = z["boxes"].shape[0], z["segs"].shape[0]
    rows = e.query_aabb(np.zeros(nq, np.int32), z["boxes"])
    for q in range(nq):
        assert np.array_equal(rows[rows[:, 0] == q, 1], z["q_fix"][z["q_off"][q]:z["q_off"][q + 1]]), q
    rows, data = e.ray_cast(np.zeros(nr, np.int32), z["segs"], closest=False)
    per = split_rows(rows, data, nr)
    crows, cdata = e.ray_cast(np.zeros(nr, np.int32), z["segs"], closest=True)
    for r in range(nr):
        a, b = z["r_off"][r], z["r_off"][r + 1]
        want_f, want_d = z["r_fix"][a:b], z["r_dat"][a:b]
        f, d = per[r]
        order = np.argsort(f, kind="stable")  # device rows come by fraction; the record by fixture
        assert np.array_equal(f[order], want_f) and np.array_equal(d[order], want_d), r
        if b > a:
            k = min(range(b - a), key=lambda i: (want_d[i, 0], want_f[i]))
            assert crows[r, 1] == want_f[k] and np.array_equal(cdata[r], want_d[k]), r
        else:
            assert crows[r, 1] == -1
    print(name, "queries", nq, "results", rows.shape[0], "rays", nr, flush=True)
    e.close()


def random_queries(rng, state, n):
    """Boxes and segments spread over each world's bodies."""
    worlds = rng.integers(0, state.n_worlds, n).astype(np.int32)
    boxes, segs = np.zeros((n, 4)), np.zeros((n, 4))
    for k, w in enumerate(worlds):
        bd = state.body_d[state.body_off[w]:state.body_off[w + 1]]
        lo, hi = bd[:, [L.BD_C_X, L.BD_C_Y]].min(0) - 2.0, bd[:, [L.BD_C_X, L.BD_C_Y]].max(0) + 2.0
        c = rng.uniform(lo, hi)
        ext = rng.uniform(0.1, 3.0, 2)
        boxes[k] = (c[0] - ext[0], c[1] - ext[1], c[0] + ext[0], c[1] + ext[1])
        segs[k, :2], segs[k, 2:] = rng.uniform(lo, hi), rng.uniform(lo, hi)
    return worlds, boxes, segs


def same_hits(f, d, wf, wd):
    """Equal fixtures, and fractions, points and normals equal to 1e-12: the reference and the oracle
    square with libm pow(x, 2) (CPython's x ** 2), which is not always correctly rounded, the
    device with x * x (correctly rounded), so the last bit can differ."""
    return np.array_equal(f, wf) and np.allclose(d, wd, rtol=1e-12, atol=1e-12)


def check_against_oracle(state, e, worlds, boxes, segs):
    o = O.Oracle(state)
    n = worlds.shape[0]
    rows = e.query_aabb(worlds, boxes)
    total = 0
    for k in range(n):
        want = o.query_aabb(int(worlds[k]), boxes[k])
        assert np.array_equal(rows[rows[:, 0] == k, 1], want), k
        total += want.shape[0]
    rrows, rdata = e.ray_cast(worlds, segs, closest=False)
    per = split_rows(rrows, rdata, n)
    crows, cdata = e.ray_cast(worlds, segs, closest=True)
    hits = 0
    for k in range(n):
        wf, wd = o.ray_cast(int(worlds[k]), segs[k])
        f, d = per[k]
        order = np.argsort(f, kind="stable")
        assert same_hits(f[order], d[order], wf, wd), k
        if f.size:
            assert crows[k, 1] == f[0] and np.array_equal(cdata[k], d[0]), k  # the first by fraction
        else:
            assert crows[k, 1] == -1, k
        hits += wf.size
    return total, hits


def test_batched_queries_over_many_worlds_match_the_oracle():
    counts = scenes.mixed_batch_counts(4096, 1)
    worlds = list(range(0, 4096, 64))
    batch = builder.scenes_batch([scenes.batch_world(1, w, int(counts[w])) for w in worlds])
    e = engine_with(batch, "fp32")
    e.step_n(200, DT, 8, 3, graph=True)
    state = e.download()
    e.close()
    e = engine_with(state)  # the oracle and the device read the same transforms and fat AABBs
    q = random_queries(np.random.default_rng(3), state, 3000)
    total, hits = check_against_oracle(state, e, *q)
    print("64 worlds: 3000 boxes ->", total, "fixtures; 3000 rays ->", hits, "hits", flush=True)
    assert total > 3000 and hits > 1000
    e.close()


@pytest.mark.parametrize("count", [200, 640])
def test_queries_on_large_worlds_match_the_oracle(count):
    """One world of 200 pyramids (11 200 proxies: the LDS sort) and of 640 (36 480 proxies: the
    global-memory sort), after 30 steps."""
    batch = builder.scene_batch(scenes.pyramids(count))
    e = engine_with(batch, "fp32")
    e.step_n(30, DT, 8, 3, graph=True)
    state = e.download()
    e.close()
    e = engine_with(state)
    total, hits = check_against_oracle(state, e, *random_queries(np.random.default_rng(4), state, 400))
    print(count, "pyramids: query results", total, "ray hits", hits, flush=True)
    assert total > 0 and hits > 0
    e.close()


def test_drop_in_world_queries_and_ray_cast_callbacks():
    import pypybox2d_amd as b2
    world = b2.World(gravity=(0, -10), do_sleep=True, continuous_physics=False)
    ground = world.create_static_body(position=(0, 0))
    ground.create_edge_fixture((-20, 0), (20, 0))
    boxes = [world.create_dynamic_body(position=(x, 0.5)) for x in (-4.0, -2.0, 0.0, 2.0, 4.0)]
    for b in boxes:
        b.create_polygon_fixture(box=(0.5, 0.5), density=1.0)
    for _ in range(10):
        world.step(DT, 8, 3)
    fixtures = [f for b in world.bodies for f in b.fixtures]
    found = list(world.query_aabb(b2.AABB(lower_bound=(-2.5, 0.2), upper_bound=(0.2, 0.8))))
    assert found == [fixtures[2], fixtures[3]]
    seen = []
    assert list(world.query_aabb(((-10, -1), (10, 2)), lambda f: seen.append(f) or len(seen) < 3)) == []
    assert len(seen) == 3
    # the generator: each yielded list's last entry clips the ray (the fraction itself, unless the
    # caller changes it), so by default only the nearest box is reported ...
    hits = list(world.ray_cast((-10.0, 0.5), (10.0, 0.5)))
    assert [h[0] for h in hits] == [fixtures[1]]
    # ... and every box, nearest first, when each hit's entry is set to 1 (continue)
    hits = []
    for ret in world.ray_cast((-10.0, 0.5), (10.0, 0.5)):
        hits.append(list(ret))
        ret[-1] = 1.0
    assert [h[0] for h in hits] == fixtures[1:]
    assert [h[3] for h in hits] == sorted(h[3] for h in hits)
    x0 = hits[0][1].x
    # the polygon ray cast tests the polygon itself, not its skin radius (shapes.py:265-307)
    assert abs(x0 - (-4.5)) < 1e-3 and abs(hits[0][2].x + 1.0) < 1e-6
    # callbacks: closest hit (clip), any hit (terminate), ignore all (-1)
    best = []
    list(world.ray_cast((-10.0, 0.5), (10.0, 0.5), lambda f, p, n, fr: best.append((f, fr)) or fr))
    assert [f for f, _ in best] == [fixtures[1]]
    first = []
    list(world.ray_cast((10.0, 0.5), (-10.0, 0.5), lambda f, p, n, fr: first.append(f) or 0.0))
    assert first == [fixtures[5]]
    every = []
    list(world.ray_cast((-10.0, 0.5), (10.0, 0.5), lambda f, p, n, fr: every.append(f) or -1.0))
    assert every == fixtures[1:]
    with pytest.raises(ValueError):
        list(world.ray_cast((1.0, 1.0), (1.0, 1.0)))


@pytest.mark.parametrize("count,steps", [(200, 40), (640, 12)])
def test_tree_pair_search_gives_the_all_pairs_contacts(count, steps):
    """The large-world pair search by LBVH traversal against the all-pairs test: bit-identical
    states (the same pairs in the same canonical order) after every step, FP32 and FP64."""
    batch = builder.scene_batch(scenes.pyramids(count))
    for prec in ("fp64", "fp32"):
        a, b = engine_with(batch, prec, tree_pairs=0), engine_with(batch, prec, tree_pairs=1)
        assert b.get_param("tree_pairs") == 1
        for s in range(steps):
            a.step(DT, 8, 3)
            b.step(DT, 8, 3)
            if s % 4 == 3 or s == steps - 1:
                sa, sb = a.download(), b.download()
                assert np.array_equal(sa.con_i, sb.con_i), (prec, s)
                assert np.array_equal(sa.body_d, sb.body_d) and np.array_equal(sa.con_d, sb.con_d), (prec, s)
        b.step_n(4, DT, 8, 3, graph=True)
        a.step_n(4, DT, 8, 3, graph=True)
        sa, sb = a.download(), b.download()
        assert np.array_equal(sa.body_d, sb.body_d) and np.array_equal(sa.con_i, sb.con_i)
        print(count, "pyramids", prec, "contacts", sa.con_i.shape[0], flush=True)
        a.close()
        b.close()
