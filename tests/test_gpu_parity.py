"""GPU parity: the HIP pipeline through the C-ABI against the FP64 oracle.

Bars:
  * broad phase: the pair list equals the oracle's exactly (set and order);
  * narrow phase, FP64: integer fields exact, floats within 1e-9;
    FP32 (FP64 geometry and evaluation, FP32 manifold storage): integer fields
    exact, floats within 2e-6;
  * one step from an identical state: relative error (floors: 1 m, 1 rad,
    |g| dt, m_min |g| dt) <= 1e-10 in FP64 and <= 1e-5 in FP32 for positions,
    angles, velocities and accumulated normal and tangent impulses; contact
    lists and every contact's discrete state identical;
  * graph replay equals eager stepping bit for bit.
"""
import os

import numpy as np
import pytest

import parity_util as PU
from conftest import GOLDEN
from oracle import oracle as O
from pypybox2d_amd import _native as N
from pypybox2d_amd import builder, layout as L, scenes

pytestmark = pytest.mark.gpu

DT = 1.0 / 60


def load_init(name):
    z = np.load(os.path.join(GOLDEN, name))
    return L.BatchArrays(*[z["init_" + n] for n in L._FIELDS]), z


def golden_state(z, init, s):
    """Full batch at checkpoint s of a recorded reference run."""
    b = init.copy()
    b.body_d, b.body_i = z["s%d_body_d" % s].copy(), z["s%d_body_i" % s].copy()
    b.fix_d = z["s%d_fix_d" % s].copy()
    b.con_d, b.con_i = z["s%d_con_d" % s].copy(), z["s%d_con_i" % s].copy()
    b.move = z["s%d_move" % s].copy()
    b.con_off = np.array([0, b.con_d.shape[0]], np.int32)
    b.move_off = np.array([0, b.move.shape[0]], np.int32)
    b.world_d[0, L.WD_INV_DT0] = 1.0 / DT
    b.world_i[0, L.WI_FLAGS] &= ~L.WF_NEW_FIXTURE
    return b


def gpu_step(batch, prec, n=1, graph=False):
    e = N.Engine(prec)
    e.upload(batch)
    if n == 1 and not graph:
        e.step(DT, 8, 3)
    else:
        e.step_n(n, DT, 8, 3, graph)
    st = e.download()
    e.close()
    return st


def oracle_step(batch, n=1):
    o = O.Oracle(batch)
    for _ in range(n):
        assert o.step(DT, 8, 3) == 0
    return o.state()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_narrowphase_seam_matches_oracle(prec):
    z = np.load(os.path.join(GOLDEN, "narrowphase.npz"))
    cd, ci = N.narrowphase(prec, z["fa_d"], z["fa_i"], z["fb_d"], z["fb_i"], z["xfa"], z["xfb"])
    n = cd.shape[0]
    print("narrowphase case count", n, prec, flush=True)
    ints_ok = np.all(ci[:, [L.CI_MTYPE, L.CI_PC, L.CI_ID0, L.CI_ID1]] ==
                     z["con_i"][:, [L.CI_MTYPE, L.CI_PC, L.CI_ID0, L.CI_ID1]], axis=1)
    pc_ok = ci[:, L.CI_PC] == z["con_i"][:, L.CI_PC]
    tol = 1e-9 if prec == "fp64" else 2e-6
    cols = [L.CD_LN_X, L.CD_LN_Y, L.CD_LP_X, L.CD_LP_Y]
    worst = 0.0
    for k in np.nonzero(ints_ok & (ci[:, L.CI_PC] > 0))[0]:
        pc = ci[k, L.CI_PC]
        c = cols + [L.CD_P0_X, L.CD_P0_Y] + ([L.CD_P1_X, L.CD_P1_Y] if pc == 2 else [])
        worst = max(worst, float(np.max(np.abs(cd[k, c] - z["con_d"][k, c]))))
    print(prec, "int mismatches", int((~ints_ok).sum()), "pc mismatches", int((~pc_ok).sum()),
          "worst float diff", worst, flush=True)
    assert ints_ok.all(), np.nonzero(~ints_ok)[0][:10]
    assert worst <= tol


def test_pair_seam_matches_oracle_exactly():
    rng = np.random.default_rng(5)
    for trial in range(6):
        n = int(rng.integers(50, 400))
        lo = rng.uniform(-20, 20, (n, 2))
        if trial % 2:  # snap to a grid so boxes touch exactly (inclusive overlap test)
            lo = np.round(lo * 2) / 2
        size = rng.uniform(0.2, 2.0, (n, 2)) if trial % 2 == 0 else np.full((n, 2), 0.5)
        fat = np.concatenate([lo, lo + size], axis=1)
        move = rng.permutation(n)[: int(rng.integers(1, n))].astype(np.int32)
        want = O.pairs(fat, move)
        got = N.pairs(fat, move)
        print("pairs trial", trial, "n", n, "moves", len(move), "pairs", len(want), flush=True)
        assert got.shape == want.shape
        assert np.array_equal(got, want)


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-10), ("fp32", 1e-5)])
@pytest.mark.parametrize("run", ["pyramid_600.npz", "stacks_600.npz", "mixed24_300.npz", "filtered_kinematic_300.npz"])
def test_one_step_from_reference_states(run, prec, tol):
    init, z = load_init(run)
    for s in [int(x) for x in z["checkpoints"]][::3]:
        b = golden_state(z, init, s)
        print("one step", run, "from checkpoint", s, prec, flush=True)
        ref = oracle_step(b)
        got = gpu_step(b, prec)
        PU.assert_one_step(got, ref, tol, (run, s))
        assert np.array_equal(got.body_i[:, :2], ref.body_i[:, :2]), (run, s)


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-10), ("fp32", 1e-5)])
def test_one_step_from_a_near_tie_state(prec, tol):
    """A recorded state in which contact (229, 290) touches with one point by less than the
    FP32 rounding of its shapes (FP32-rounded shapes give no point): both pipelines must
    reproduce the oracle's discrete state for every contact."""
    z = np.load(os.path.join(GOLDEN, "near_tie_mixed3975.npz"))
    b = L.BatchArrays(*[z[k] for k in L._FIELDS])
    b.world_i[:, L.WI_FLAGS] &= ~L.WF_NEW_FIXTURE
    print("near-tie state", prec, "bodies", b.body_d.shape[0], "contacts", b.con_d.shape[0], flush=True)
    ref = oracle_step(b)
    got = gpu_step(b, prec)
    assert PU.same_contact_lists(got, ref)
    cols = [L.CI_FA, L.CI_FB, L.CI_FLAGS, L.CI_MTYPE, L.CI_PC, L.CI_ID0, L.CI_ID1]
    assert np.array_equal(got.con_i[:, cols], ref.con_i[:, cols])
    k = np.nonzero((ref.con_i[:, L.CI_FA] == 229) & (ref.con_i[:, L.CI_FB] == 290))[0][0]
    assert ref.con_i[k, L.CI_PC] == 1
    PU.assert_one_step(got, ref, tol, "near tie")


def test_first_step_from_construction_fp64():
    """Start from the freshly built world (new-fixture pair search on the device)."""
    for run in ["pyramid_600.npz", "stacks_600.npz", "mixed24_300.npz", "filtered_kinematic_300.npz"]:
        init, z = load_init(run)
        ref = oracle_step(init, 3)
        got = gpu_step(init, "fp64", 3)
        assert PU.same_contact_lists(got, ref), run
        be = PU.body_errors(got, ref)
        print(run, "3 steps from construction", be, flush=True)
        assert max(be.values()) <= 1e-10


def test_pyramid_trajectory_fp64_tracks_oracle():
    init, z = load_init("pyramid_600.npz")
    o = O.Oracle(init)
    e = N.Engine("fp64")
    e.upload(init)
    for s in range(1, 601):
        assert o.step(DT, 8, 3) == 0
        e.step(DT, 8, 3)
        if s in (10, 100, 300, 600):
            ref, got = o.state(), e.download()
            be = PU.body_errors(got, ref)
            print("pyramid fp64 step", s, be, flush=True)
            assert be["position"] <= 1e-6, (s, be)
    e.close()


@pytest.mark.parametrize("large_worlds", [0, 1])
def test_filtered_kinematic_trajectory_fp64_tracks_oracle(large_worlds):
    """The recorded run with category/mask/group filtering and a kinematic platform, 300 steps
    from construction on the FP64 pipeline (LDS searches, then the global-memory searches),
    against the oracle: identical contact lists at every checkpoint (the filter decides which
    overlapping pairs become contacts), body states within 1e-12 for the first 100 steps and
    within 1e-6 after (the scene tumbles; last-bit differences of the device's sin/cos grow)."""
    init, z = load_init("filtered_kinematic_300.npz")
    o = O.Oracle(init)
    e = N.Engine("fp64")
    e.set_param("large_worlds", large_worlds)
    e.upload(init)
    for s in range(1, 301):
        assert o.step(DT, 8, 3) == 0
        e.step(DT, 8, 3)
        if s in (1, 20, 100, 200, 300):
            ref, got = o.state(), e.download()
            assert PU.same_contact_lists(got, ref), s
            be = PU.body_errors(got, ref)
            print("filtered/kinematic fp64 step", s, "large_worlds", large_worlds, be, flush=True)
            assert max(be[q] for q in PU.QUANTITIES) <= (1e-12 if s <= 100 else 1e-6), (s, be)
    e.close()


@pytest.mark.parametrize("large_worlds", [0, 1])
def test_a_touched_proxy_is_searched_at_the_end_of_the_step(large_worlds):
    """A move-buffer entry the batch brings without new fixtures (Fixture.refilter touching a proxy,
    fixture.py:282-285) stays in the buffer until the end-of-step pair search, ahead of the proxies
    the synchronisation moves (broadphase.py:61-68, 86-99); a dt = 0 step keeps it (no
    find_new_contacts runs).  From the recorded filtered run at step 150, the first category-2
    fixture becomes category 1 with a full mask, its contacts are flagged for filtering and its
    proxy touched: the pairs it gains (two, which the untouched buffer would miss) must be created
    as the oracle creates them."""
    init, z = load_init("filtered_kinematic_300.npz")

    def edited(touch):
        b = golden_state(z, init, 150)
        f = [k for k in range(b.fix_i.shape[0]) if b.fix_i[k, L.FI_CATEGORY] == 2][0]
        b.fix_i[f, L.FI_CATEGORY], b.fix_i[f, L.FI_MASK] = 1, 0xFFFF
        for k in range(b.con_i.shape[0]):
            if f in (b.con_i[k, L.CI_FA], b.con_i[k, L.CI_FB]):
                b.con_i[k, L.CI_FLAGS] |= L.CF_FILTER
        if touch:
            b.move = np.array([f], np.int32)
            b.move_off = np.array([0, 1], np.int32)
        return b, f
    b, f = edited(True)
    untouched = O.Oracle(edited(False)[0])
    untouched.step(DT, 8, 3)
    o = O.Oracle(b)
    e = N.Engine("fp64")
    e.set_param("large_worlds", large_worlds)
    e.upload(b)
    o.step(0.0, 8, 3)
    e.step(0.0, 8, 3)
    ref, got = o.state(), e.download()
    assert list(got.move) == [f] and list(ref.move) == [f]  # a dt = 0 step keeps the buffer
    for s in range(1, 31):
        assert o.step(DT, 8, 3) == 0
        e.step(DT, 8, 3)
        ref, got = o.state(), e.download()
        assert PU.same_contact_lists(got, ref), s
        if s == 1:
            assert ref.con_i.shape[0] > untouched.state().con_i.shape[0]
    be = PU.body_errors(got, ref)
    assert max(be[q] for q in PU.QUANTITIES) <= 1e-12, be
    e.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_batched_mixed_worlds_one_step(prec):
    rng = np.random.default_rng(3)
    worlds = [scenes.mixed_world(rng, int(n), half_width=10.0, height=15.0) for n in (20, 45, 33, 60, 12, 80)]
    batch = builder.scenes_batch(worlds)
    settled = oracle_step(batch, 40)  # falling and colliding by now
    settled.world_i[:, L.WI_FLAGS] &= ~L.WF_NEW_FIXTURE
    ref = oracle_step(settled)
    got = gpu_step(settled, prec)
    PU.assert_one_step(got, ref, 1e-10 if prec == "fp64" else 1e-5, (prec, "mixed batch"))


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_lognormal_batch_worlds_one_step(prec):
    """Worlds of the lognormal batch (16 384 worlds, body counts clip(round(64 exp(0.75 N)),
    16, 1024), seed 2): its first 40 worlds and its largest, stepped 200 times on the GPU so
    that piles form, then one step from that state against the oracle."""
    counts = scenes.lognormal_batch_counts()
    pick = list(range(40)) + [int(np.argmax(counts))]
    batch = builder.scenes_batch([scenes.batch_world(2, w, counts[w]) for w in pick])
    assert counts[pick[-1]] == 1024
    e = N.Engine(prec)
    e.upload(batch)
    e.step_n(200, DT, 8, 3, graph=True)
    state = e.download()
    e.close()
    state.world_i[:, L.WI_FLAGS] &= ~L.WF_NEW_FIXTURE
    ref = oracle_step(state)
    got = gpu_step(state, prec)
    PU.assert_one_step(got, ref, 1e-10 if prec == "fp64" else 1e-5, (prec, "lognormal worlds"))


def settled_mixed_sample(steps=1000, n=64):
    """The worlds that bench.py samples from the 4096-world mixed batch (seed 1), stepped `steps`
    times on the GPU's FP32 pipeline (graph-captured) so that piles have formed and settled."""
    counts = scenes.mixed_batch_counts(4096, 1)
    worlds = sorted(np.random.default_rng(0).choice(4096, size=n, replace=False).tolist())
    batch = builder.scenes_batch([scenes.batch_world(1, w, int(counts[w])) for w in worlds])
    e = N.Engine("fp32")
    e.upload(batch)
    e.step_n(steps, DT, 8, 3, graph=True)
    state = e.download()
    e.close()
    state.world_i[:, L.WI_FLAGS] &= ~L.WF_NEW_FIXTURE
    return state


def test_settled_mixed_batch_sample_one_step():
    """64 worlds of the mixed batch after 1000 steps (about 18 000 bodies, most of them in piles):
    one step of each pipeline from that state against the oracle. Both pipelines keep the
    oracle's pair lists and discrete contact states; FP32 holds every body quantity and every
    accumulated normal and tangent impulse within 1e-5 relative, FP64 within 1e-10."""
    state = settled_mixed_sample()
    touching = int(np.count_nonzero(state.con_i[:, L.CI_FLAGS] & L.CF_TOUCHING))
    print("settled sample: bodies", state.body_d.shape[0], "contacts", state.con_d.shape[0], "touching", touching,
          flush=True)
    assert touching > state.body_d.shape[0]  # piles: more touching contacts than bodies
    ref = oracle_step(state)
    for prec, tol in (("fp64", 1e-10), ("fp32", 1e-5)):
        got = gpu_step(state, prec)
        PU.assert_one_step(got, ref, tol, (prec, "settled mixed batch"))


def test_ragged_batch_with_empty_and_fixtureless_worlds():
    """Worlds with no dynamic body (ground only), one body, 40 bodies, and one whose bodies
    have no fixtures (the dynamic one falls with the default unit mass): FP64 tracks the
    oracle for 60 steps with equal contact lists, then one FP32 step from that state."""
    rng = np.random.default_rng(5)
    ground_only = scenes.mixed_world(rng, 0, half_width=10.0, height=15.0)
    one = scenes.mixed_world(rng, 1, half_width=10.0, height=15.0)
    forty = scenes.mixed_world(rng, 40, half_width=10.0, height=15.0)
    bare = {"gravity": (0.0, -10.0), "do_sleep": True, "warm_starting": True,
            "bodies": [{"type": "dynamic", "position": (0.0, 3.0), "angle": 0.0, "linear_velocity": (1.0, 0.0),
                        "angular_velocity": 0.5, "fixtures": []},
                       {"type": "static", "position": (0.0, 0.0), "angle": 0.0, "linear_velocity": (0.0, 0.0),
                        "angular_velocity": 0.0, "fixtures": []}]}
    batch = builder.scenes_batch([ground_only, one, forty, bare, ground_only])
    assert list(np.diff(batch.fix_off)) == [1, 2, 41, 0, 1]
    o = O.Oracle(batch)
    e = N.Engine("fp64")
    e.upload(batch)
    for _ in range(60):
        assert o.step(DT, 8, 3) == 0
        e.step(DT, 8, 3)
    got, ref = e.download(), o.state()
    e.close()
    assert PU.same_contact_lists(got, ref)
    be = PU.body_errors(got, ref)
    print("ragged fp64, 60 steps", be, flush=True)
    assert max(be[q] for q in ("position", "angle", "velocity", "angular_velocity")) <= 1e-9
    ref.world_i[:, L.WI_FLAGS] &= ~L.WF_NEW_FIXTURE
    got32, ref1 = gpu_step(ref, "fp32"), oracle_step(ref)
    PU.assert_one_step(got32, ref1, 1e-5, "ragged fp32")


def test_graph_replay_matches_eager():
    rng = np.random.default_rng(4)
    batch = builder.scenes_batch([scenes.mixed_world(rng, n, half_width=10.0, height=15.0) for n in (30, 50, 40)])
    a, b = N.Engine("fp32"), N.Engine("fp32")
    a.upload(batch)
    b.upload(batch)
    a.step(DT, 8, 3)
    b.step(DT, 8, 3)  # the new-fixture step runs eagerly in both
    for r in range(5):
        for _ in range(4):
            a.step(DT, 8, 3)
        b.step_n(4, DT, 8, 3, graph=True)
        sa, sb = a.download(), b.download()
        print("graph replay round", r, flush=True)
        assert np.array_equal(sa.body_d, sb.body_d)
        assert np.array_equal(sa.con_d, sb.con_d) and np.array_equal(sa.con_i, sb.con_i)
    a.close()
    b.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_island_launch_shapes_agree_bit_exactly(prec):
    """The wavefront-per-island path, the lane-per-island path and the LDS
    fallback run the same solves in the same order: identical results."""
    counts = [420, 300, 512, 180, 96, 260]
    batch = builder.scenes_batch([scenes.batch_world(9, w, n) for w, n in enumerate(counts)])
    e = N.Engine(prec)
    e.upload(batch)
    e.step_n(150, DT, 8, 3, graph=True)  # piles have formed
    state = e.download()
    e.close()
    state.world_i[:, L.WI_FLAGS] &= ~L.WF_NEW_FIXTURE
    results = {}
    for name, params in (("lanes", {"heavy_class": 16}), ("teams64", {"heavy_class": 0, "team_full_class": 0}),
                         ("teams8", {"heavy_class": 0, "team_full_class": 16}), ("default", {}),
                         ("fallback", {"heavy_class": 3, "team_lds": 4096}), ("large_worlds", {"large_worlds": 1}),
                         ("pairs", {"heavy_class": 5, "team_full_class": 9, "team_max": 32, "team_lds": 24576}),
                         ("whole_wave", {"heavy_class": 3, "team_full_class": 9, "team_max": 16, "team_lds": 8192}),
                         ("narrow_searches", {"collide_block": 64, "islands_wide": 0, "pairs_wide": 0}),
                         ("wide_searches", {"collide_block": 512, "islands_wide": 32, "pairs_wide": 32})):
        e = N.Engine(prec)
        for k, v in params.items():
            e.set_param(k, v)
        e.upload(state)
        for _ in range(3):
            e.step(DT, 8, 3)
        e.step_n(3, DT, 8, 3, graph=True)
        results[name] = e.download()
        print(prec, name, "done", flush=True)
        e.close()
    ref = results["lanes"]
    for name, st in results.items():
        assert np.array_equal(st.body_d, ref.body_d), name
        assert np.array_equal(st.body_i, ref.body_i), name
        assert np.array_equal(st.con_d, ref.con_d) and np.array_equal(st.con_i, ref.con_i), name


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_worlds_of_many_sizes_launch_in_runs(prec):
    """Worlds of 13 to 1021 bodies: the island and the pair search each launch in four runs
    (worlds grouped by LDS footprint, side by side on separate streams). One step tracks the
    oracle, and the runs, eager or graph-captured, give bit for bit what the global-memory
    search gives."""
    counts = [1020, 40, 700, 300, 520, 60, 900, 12]
    batch = builder.scenes_batch([scenes.batch_world(11, w, n) for w, n in enumerate(counts)])
    e = N.Engine(prec)
    e.upload(batch)
    assert e.get_param("search_runs") == 4 | 4 << 8
    e.step_n(120, DT, 8, 3, graph=True)
    state = e.download()
    e.close()
    state.world_i[:, L.WI_FLAGS] &= ~L.WF_NEW_FIXTURE
    ref = oracle_step(state)
    got = gpu_step(state, prec)
    PU.assert_one_step(got, ref, 1e-10 if prec == "fp64" else 1e-5, (prec, "worlds in runs"))
    results = {}
    for name, params, runs in (("runs", {}, 4 | 4 << 8), ("large_worlds", {"large_worlds": 1}, 1 | 1 << 8)):
        e = N.Engine(prec)
        for k, v in params.items():
            e.set_param(k, v)
        e.upload(state)
        assert e.get_param("search_runs") == runs
        for _ in range(3):
            e.step(DT, 8, 3)
        e.step_n(3, DT, 8, 3, graph=True)
        results[name] = e.download()
        print(prec, name, "done", flush=True)
        e.close()
    a, b = results["runs"], results["large_worlds"]
    assert np.array_equal(a.body_d, b.body_d) and np.array_equal(a.body_i, b.body_i)
    assert np.array_equal(a.con_d, b.con_d) and np.array_equal(a.con_i, b.con_i)


def test_set_param_rejects_bad_values():
    e = N.Engine("fp32")
    assert e.get_param("heavy_class") >= 0
    for k, v in (("heavy_class", 17), ("team_lds", 200000), ("team_lds", 1000), ("solve_grid", 0),
                 ("team_full_class", -1), ("team_max", 24), ("team_max", 4), ("team_max", 128),
                 ("search_runs", 1), ("solve_waves_per_cu", 8), ("collide_block", 96), ("collide_block", 1024),
                 ("collide_block", 32), ("pairs_wide", 33), ("islands_wide", -1)):
        with pytest.raises(N.B2GError):
            e.set_param(k, v)
    e.close()


def test_restitution_trajectory_fp64_and_one_step_fp32():
    """Bouncing bodies (restitution 0.6) exercise the restitution bias of the velocity constraints."""
    rng = np.random.default_rng(11)
    sc = scenes.mixed_world(rng, 40, half_width=8.0, height=12.0)
    for b in sc["bodies"]:
        for f in b["fixtures"]:
            f["restitution"] = 0.6
    batch = builder.scene_batch(sc)
    o = O.Oracle(batch)
    e = N.Engine("fp64")
    e.upload(batch)
    for s in range(60):
        o.step(DT, 8, 3)
        e.step(DT, 8, 3)
    got, ref = e.download(), o.state()
    e.close()
    assert PU.same_contact_lists(got, ref)
    be = PU.body_errors(got, ref)
    print("restitution fp64 60 steps", be, flush=True)
    assert max(be[q] for q in ("position", "angle", "velocity", "angular_velocity")) <= 1e-9
    ref.world_i[:, L.WI_FLAGS] &= ~L.WF_NEW_FIXTURE
    one = oracle_step(ref)
    g32 = gpu_step(ref, "fp32")
    PU.assert_one_step(g32, one, 1e-5, "restitution fp32")


def test_drop_in_world_api_tracks_oracle():
    """The pypybox2d-style World/Body API, stepped on the GPU, against the oracle on the same scene."""
    import pypybox2d_amd as b2
    world = b2.World(gravity=(0, -10), do_sleep=True, continuous_physics=False)
    ground = world.create_static_body(position=(0, 0))
    ground.create_edge_fixture((-40, 0), (40, 0))
    boxes = []
    for i in range(10):
        for j in range(10 - i):
            b = world.create_dynamic_body(position=(-7 + 0.5625 * i + 1.125 * j, 0.75 + 1.0 * i))
            b.create_polygon_fixture(box=(0.5, 0.5), density=5, friction=0.6)
            boxes.append(b)
    o = O.Oracle(builder.scene_batch(scenes.pyramid()))
    for s in range(120):
        world.step(DT, 8, 3)
        o.step(DT, 8, 3)
        if s % 40 == 39:
            ref = o.state().body_d
            worst = 0.0
            for k, b in enumerate(boxes):  # body 0 of the scene is the ground
                p, r = b.position, ref[k + 1]
                worst = max(worst, abs(p.x - r[L.BD_XF_PX]), abs(p.y - r[L.BD_XF_PY]),
                            abs(b.angle - r[L.BD_A]), abs(b.linear_velocity.y - r[L.BD_V_Y]))
            print("drop-in world step", s + 1, "max abs diff", worst, flush=True)
            assert worst <= 1e-9
    assert world.contact_count == o.state().con_i.shape[0]


def test_stacks_2000_steps_settle_height():
    """Vertical stacks, 2000 steps: the FP64 pipeline tracks the oracle; the FP32 pipeline's
    settle height (mean top-box height over the last 100 steps) stays within 1 mm of it."""
    batch = builder.scene_batch(scenes.stacks(seed=0))
    o = O.Oracle(batch)
    e64, e32 = N.Engine("fp64"), N.Engine("fp32")
    e64.upload(batch)
    e32.upload(batch)
    tops = [k for k in range(1, 51) if k % 10 == 0]  # the top box of each column
    h = {"ref": [], "fp64": [], "fp32": []}
    for s in range(2000):
        o.step(DT, 8, 3)
        e64.step(DT, 8, 3)
        e32.step(DT, 8, 3)
        if s >= 1900:
            for name, st in (("ref", o.state()), ("fp64", e64.download()), ("fp32", e32.download())):
                h[name].append(st.body_d[tops, L.BD_C_Y].mean())
    ref, g64, g32 = o.state(), e64.download(), e32.download()
    e64.close()
    e32.close()
    settle = {k: float(np.mean(v)) for k, v in h.items()}
    be = PU.body_errors(g64, ref)
    print("stacks 2000: settle heights", settle, "fp64 errors", be, flush=True)
    assert max(be[q] for q in ("position", "angle")) <= 1e-9
    assert abs(settle["fp64"] - settle["ref"]) <= 1e-9
    assert abs(settle["fp32"] - settle["ref"]) <= 1e-3


def test_large_world_path_tracks_oracle():
    """The global-memory island and pair search (forced for every world) over the pyramid's
    600 steps, through contact creation, sleep and the awake gating, against the oracle."""
    batch = builder.scene_batch(scenes.pyramid())
    o = O.Oracle(batch)
    e = N.Engine("fp64")
    e.set_param("large_worlds", 1)
    e.upload(batch)
    asleep_steps = 0
    for s in range(600):
        o.step(DT, 8, 3)
        e.step(DT, 8, 3)
        ref = o.state()
        asleep_steps += int(not np.all(ref.body_i[1:, L.BI_FLAGS] & L.BF_AWAKE))  # next step runs the gating
        if s % 100 == 99:
            got = e.download()
            assert PU.same_contact_lists(got, ref), s
            be = PU.body_errors(got, ref)
            print("large-world path step", s + 1, be, flush=True)
            assert be["position"] <= 1e-6, (s, be)
    e.close()
    print("steps that ended with a sleeping body:", asleep_steps, flush=True)
    assert asleep_steps > 0


def test_single_large_world_200_pyramids():
    """One world of 200 pyramids (11 000 boxes, 200 ground edges; 11 200 proxies): too large for
    the LDS kernels, so it runs on the global-memory path. FP64 tracks the oracle for 60 steps;
    FP32 one step from there is within 1e-5."""
    batch = builder.scene_batch(scenes.pyramids(200))
    o = O.Oracle(batch)
    e = N.Engine("fp64")
    e.upload(batch)
    for s in range(60):
        o.step(DT, 8, 3)
        e.step(DT, 8, 3)
    got, ref = e.download(), o.state()
    e.close()
    assert PU.same_contact_lists(got, ref)
    be = PU.body_errors(got, ref)
    print("200 pyramids fp64, 60 steps", be, flush=True)
    assert max(be[q] for q in ("position", "angle", "velocity", "angular_velocity")) <= 1e-9
    ref.world_i[:, L.WI_FLAGS] &= ~L.WF_NEW_FIXTURE
    one = oracle_step(ref)
    g32 = gpu_step(ref, "fp32")
    PU.assert_one_step(g32, one, 1e-5, "200 pyramids fp32")


def test_single_world_with_body_indices_above_32767():
    """One world of 640 pyramids (35 840 bodies, about 64 000 contacts): the large-world path's
    16-bit union-find parents and packed fixture/body indices run above 32767. FP64 tracks the
    oracle for 20 steps."""
    batch = builder.scene_batch(scenes.pyramids(640))
    assert batch.body_off[1] > 32768
    o = O.Oracle(batch)
    e = N.Engine("fp64")
    e.upload(batch)
    for s in range(20):
        o.step(DT, 8, 3)
        e.step(DT, 8, 3)
    got, ref = e.download(), o.state()
    e.close()
    assert PU.same_contact_lists(got, ref)
    be = PU.body_errors(got, ref)
    print("640 pyramids fp64, 20 steps", be, flush=True)
    assert max(be[q] for q in ("position", "angle", "velocity", "angular_velocity")) <= 1e-9


def test_island_times_report_every_island_and_change_nothing():
    counts = [420, 300, 512, 180, 96, 260]
    batch = builder.scenes_batch([scenes.batch_world(9, w, n) for w, n in enumerate(counts)])
    out = {}
    for timed in (0, 1):
        e = N.Engine("fp32")
        e.set_param("island_times", timed)
        e.upload(batch)
        e.step_n(60, DT, 8, 3, graph=True)
        out[timed] = e.download()
        if timed:
            rows = e.island_times()
            waves = e.get_param("solve_waves_per_cu")
            with pytest.raises(N.B2GError):
                e.set_param("solve_waves_per_cu", 4)
        e.close()
    assert np.array_equal(out[0].body_d, out[1].body_d)
    assert np.array_equal(out[0].con_d, out[1].con_d)
    print("islands", rows.shape[0], "waves per CU", waves, "lanes", sorted(set(rows[:, 4].tolist())), flush=True)
    assert rows.shape[0] > 0 and 1 <= waves <= 32
    assert np.all(rows[:, 1] >= rows[:, 0])
    assert set(rows[:, 4].tolist()) <= {0, 1, 8, 16, 32, 64}
    assert rows.shape[1] == 8 and np.all(rows[rows[:, 4] >= 8, 5] >= 1)  # team islands' sweep periods
    assert np.all((rows[:, 6] >= rows[:, 0]) & (rows[:, 6] <= rows[:, 1]))  # the velocity phase ends inside
    assert np.all((rows[:, 7] >= 1) & (rows[:, 7] <= 3))  # position iterations run (early exit)
    dyn = int(np.count_nonzero(batch.body_i[:, L.BI_TYPE] == L.DYNAMIC))
    assert rows[:, 3].sum() <= dyn and np.all(rows[:, 3] >= 1)
