"""GPU: error paths, BatchedWorld and the launch tuner, through the C-ABI.

  * a batch with too few contact slots reports B2G_E_CAPACITY at the next sync
    (the kernels flag the world and end normally);
  * World and BatchedWorld grow the slots and replay from the last good
    state, bit for bit what a batch with enough slots gives;
  * a non-finite body state reports B2G_E_NAN;
  * BatchedWorld.step(n, graph=True) with the launch tuner equals eager
    Engine steps bit for bit, and the tuner's choices are launch shapes of
    its candidate list.
"""
import numpy as np
import pytest

from pypybox2d_amd import _native as N
from pypybox2d_amd import builder, layout as L, scenes
from pypybox2d_amd.autotune import LaunchTuner
from pypybox2d_amd.batched import BatchedWorld

pytestmark = pytest.mark.gpu

DT = 1.0 / 60


def pyramid_world(**kw):
    import pypybox2d_amd as b2
    world = b2.World(gravity=(0, -10), do_sleep=True, continuous_physics=False, **kw)
    ground = world.create_static_body(position=(0, 0))
    ground.create_edge_fixture((-40, 0), (40, 0))
    boxes = []
    for i in range(10):
        for j in range(10 - i):
            b = world.create_dynamic_body(position=(-7 + 0.5625 * i + 1.125 * j, 0.75 + 1.0 * i))
            b.create_polygon_fixture(box=(0.5, 0.5), density=5, friction=0.6)
            boxes.append(b)
    return world, boxes


def test_too_few_contact_slots_report_a_capacity_error():
    batch = builder.scene_batch(scenes.pyramid())
    e = N.Engine("fp32", contact_factor=0.01, contact_extra=1)
    e.upload(batch)
    print("capacity: stepping a 56-fixture pyramid with 16 contact slots", flush=True)
    e.step(DT, 8, 3)  # the first step's pair search finds about 150 pairs
    with pytest.raises(N.B2GError) as exc:
        e.sync()
    print("reported:", exc.value, flush=True)
    assert exc.value.code == N.B2G_E_CAPACITY
    e.close()


def test_world_regrows_contact_slots_and_matches_a_roomy_world():
    small, boxes_s = pyramid_world(contact_factor=0.01)
    roomy, boxes_r = pyramid_world()
    for s in range(40):
        small.step(DT, 8, 3)
        roomy.step(DT, 8, 3)
    print("regrows", small._stepper.regrows, "slots per fixture", small._stepper.contact_factor, flush=True)
    assert small._stepper.regrows >= 1
    for a, b in zip(boxes_s, boxes_r):
        assert (a.position.x, a.position.y, a.angle) == (b.position.x, b.position.y, b.angle)
        assert (a.linear_velocity.x, a.linear_velocity.y) == (b.linear_velocity.x, b.linear_velocity.y)
    assert small.contact_count == roomy.contact_count


def test_non_finite_state_is_reported():
    sc = scenes.mixed_world(np.random.default_rng(2), 12, half_width=8.0, height=10.0)
    batch = builder.scene_batch(sc)
    k = int(np.nonzero(batch.body_i[:, L.BI_TYPE] == L.DYNAMIC)[0][3])
    batch.body_d[k, L.BD_V_X] = np.nan
    e = N.Engine("fp64")
    e.upload(batch)
    print("non-finite: one body with a NaN velocity, one step", flush=True)
    e.step(DT, 8, 3)
    with pytest.raises(N.B2GError) as exc:
        e.sync()
    print("reported:", exc.value, flush=True)
    assert exc.value.code == N.B2G_E_NAN
    e.close()


def mixed_batch(seed, counts):
    return builder.scenes_batch([scenes.batch_world(seed, w, n) for w, n in enumerate(counts)])


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_batched_world_graph_and_tuner_equal_eager_steps(prec):
    batch = mixed_batch(13, [300, 120, 480, 64, 250, 33])
    bw = BatchedWorld(batch, prec, autotune=True)
    bw.tuner.early = [10, 10]  # tune three times within the run
    bw.tuner.period = 15
    bw.step(DT, 8, 3, n=45)
    got = bw.state()
    print(prec, "tuned launch shape", bw.tuner.best, "tunes", bw.tuner.tunes, flush=True)
    assert bw.tuner.tunes >= 3
    for name, _, cands in bw.tuner.groups:
        assert bw.tuner.choice[name] in cands
    e = N.Engine(prec)
    e.upload(batch)
    for _ in range(45):
        e.step(DT, 8, 3)
    ref = e.download()
    e.close()
    bw.close()
    assert np.array_equal(got.body_d, ref.body_d) and np.array_equal(got.body_i, ref.body_i)
    assert np.array_equal(got.con_d, ref.con_d) and np.array_equal(got.con_i, ref.con_i)


def test_batched_world_regrows_contact_slots_bit_exactly():
    batch = mixed_batch(14, [200, 90, 310])
    bw = BatchedWorld(batch, "fp32", contact_factor=0.05, contact_extra=4, autotune=False)
    bw.step(DT, 8, 3, n=30)
    got = bw.state()
    print("batched regrows", bw.stepper.regrows, "slots per fixture", bw.stepper.contact_factor, flush=True)
    assert bw.stepper.regrows >= 1
    bw.step(DT, 8, 3, n=10)  # and on from the downloaded checkpoint
    got2 = bw.state()
    bw.close()
    e = N.Engine("fp32")
    e.upload(batch)
    e.step_n(30, DT, 8, 3, graph=True)
    ref = e.download()
    e.step_n(10, DT, 8, 3, graph=False)
    ref2 = e.download()
    e.close()
    for a, b in ((got, ref), (got2, ref2)):
        assert np.array_equal(a.body_d, b.body_d) and np.array_equal(a.con_i, b.con_i)
        assert np.array_equal(a.con_d, b.con_d)


def test_launch_tuner_picks_the_fastest_candidate_it_timed():
    batch = mixed_batch(15, [400, 380, 512, 200])
    e = N.Engine("fp32")
    e.upload(batch)
    e.step_n(60, DT, 8, 3, graph=True)
    t = LaunchTuner(e)
    best = t.tune(DT, 8, 3)
    print("candidate solve ms", np.round(t.last_times, 3).tolist(), "best", best, flush=True)
    assert t.last_ms.shape == (t.steps_per_tune, len(N.STAGES))
    assert t.choice["solve"] == t.candidates[int(np.argmin(t.last_times))]
    for name, stage, cands in t.groups:  # each group: the candidate of least mean stage time
        runs = [[i for i in range(t.steps_per_tune) if i % len(cands) == c] for c in range(len(cands))]
        mean = [t.last_ms[r, stage].mean() for r in runs]
        assert t.choice[name] == cands[int(np.argmin(mean))], name
    assert set(best) == {"heavy_class", "team_full_class", "team_max", "team_lds", "collide_block",
                         "islands_wide", "pairs_wide"}
    for k, v in best.items():
        assert e.get_param(k) == v
    e.close()
