"""Closed forms and invariants that pin the FP64 oracle beyond the recorded runs.

Each check is derived from the mathematics of the step,
not from the oracle's own formulas: free flight (semi-implicit Euler,
island.py:71-138), the restitution bias (contact.py:451-456) exchanging the
velocities of equal head-on circles, the friction cone and non-negative normal
impulses (contact.py:548-596), the symmetric block solution of a box resting
flat (contact.py:600-640), and the pair set against brute force
(dynamictree_test.py:89-98 tests the reference's tree the same way).
"""
import numpy as np
import pytest

from oracle import oracle as O
from pypybox2d_amd import builder, layout as L, scenes

DT = 1.0 / 60


def _scene(bodies, gravity=(0.0, -10.0)):
    return {"gravity": gravity, "do_sleep": False, "warm_starting": True, "bodies": bodies}


def _body(t, pos, shape, v=(0.0, 0.0), w=0.0, angle=0.0, density=1.0, friction=0.3, restitution=0.0):
    return {"type": t, "position": pos, "angle": angle, "linear_velocity": v, "angular_velocity": w,
            "fixtures": [{"shape": shape, "density": density, "friction": friction,
                          "restitution": restitution}]}


def test_free_flight_is_semi_implicit_euler():
    g = (0.0, -10.0)
    v0, w0 = (1.5, 2.0), 0.7
    st = builder.scene_batch(_scene([_body("dynamic", (0.0, 50.0), ("circle", 0.5), v=v0, w=w0)], g))
    o = O.Oracle(st)
    o.step(DT, 8, 3)
    s = o.state().body_d[0]
    vx, vy = v0[0] + DT * g[0], v0[1] + DT * g[1]
    assert s[L.BD_V_X] == pytest.approx(vx, rel=1e-15, abs=1e-15)
    assert s[L.BD_V_Y] == pytest.approx(vy, rel=1e-15)
    assert s[L.BD_C_X] == pytest.approx(0.0 + DT * vx, rel=1e-14)
    assert s[L.BD_C_Y] == pytest.approx(50.0 + DT * vy, rel=1e-15)
    assert s[L.BD_A] == pytest.approx(DT * w0, rel=1e-15)
    assert s[L.BD_W] == w0


def test_translation_clamp():
    st = builder.scene_batch(_scene([_body("dynamic", (0.0, 50.0), ("circle", 0.5), v=(600.0, 0.0))], (0.0, 0.0)))
    o = O.Oracle(st)
    o.step(DT, 8, 3)
    s = o.state().body_d[0]
    assert s[L.BD_C_X] == pytest.approx(2.0, rel=1e-14)  # MAX_TRANSLATION per step
    assert s[L.BD_V_X] == pytest.approx(120.0, rel=1e-14)


@pytest.mark.parametrize("e", [1.0, 0.5])
def test_head_on_circles_restitution(e):
    """Equal circles approaching at 4 m/s (> the 1 m/s threshold): v_n' = -e v_n."""
    bodies = [_body("dynamic", (-0.49, 5.0), ("circle", 0.5), v=(2.0, 0.0), restitution=e, friction=0.0),
              _body("dynamic", (0.49, 5.0), ("circle", 0.5), v=(-2.0, 0.0), restitution=e, friction=0.0)]
    o = O.Oracle(builder.scene_batch(_scene(bodies, (0.0, 0.0))))
    o.step(DT, 8, 3)
    s = o.state()
    va, vb = s.body_d[0, L.BD_V_X], s.body_d[1, L.BD_V_X]
    assert vb - va == pytest.approx(4.0 * e, rel=1e-12)  # relative normal velocity reversed and scaled
    assert va + vb == pytest.approx(0.0, abs=1e-12)  # momentum
    assert s.con_i[0, L.CI_PC] == 1


def test_box_resting_flat_has_symmetric_block_solution():
    """A box centred over a static edge, first step (no warm start): case 1 of the block
    solver with a symmetric K and b gives equal normal impulses."""
    bodies = [_body("static", (0.0, 0.0), ("edge", (-5.0, 0.0), (5.0, 0.0))),
              _body("dynamic", (0.0, 0.499), ("box", 0.5, 0.5), density=2.0)]
    o = O.Oracle(builder.scene_batch(_scene(bodies)))
    o.step(DT, 8, 3)
    s = o.state()
    assert s.con_i[0, L.CI_PC] == 2
    n0, n1 = s.con_d[0, L.CD_P0_NI], s.con_d[0, L.CD_P1_NI]
    assert n0 > 0 and n1 > 0
    assert n0 == pytest.approx(n1, rel=1e-12)
    # the pair together cancels one step of gravity (plus the restitution-free approach)
    m = s.body_d[1, L.BD_MASS]
    assert n0 + n1 >= m * 10.0 * DT * (1 - 1e-12)


def test_friction_cone_and_nonnegative_normal_impulses():
    st = builder.scene_batch(scenes.stacks(seed=3))
    o = O.Oracle(st)
    worst = 0.0
    for k in range(120):
        o.step(DT, 8, 3)
        if k % 10 == 9:
            s = o.state()
            live = (s.con_i[:, L.CI_FLAGS] & L.CF_TOUCHING) != 0
            for p, (ni, ti) in enumerate(((L.CD_P0_NI, L.CD_P0_TI), (L.CD_P1_NI, L.CD_P1_TI))):
                m = live & (s.con_i[:, L.CI_PC] > p)
                lam_n, lam_t, mu = s.con_d[m, ni], s.con_d[m, ti], s.con_d[m, L.CD_FRICTION]
                assert np.all(lam_n >= 0.0)
                worst = max(worst, float(np.max(np.abs(lam_t) - mu * lam_n, initial=-1.0)))
    assert worst <= 1e-12


def _brute_pairs(fat, move):
    def overlap(a, b):
        return not (b[0] - a[2] > 0 or b[1] - a[3] > 0 or a[0] - b[2] > 0 or a[1] - b[3] > 0)
    out = set()
    for i in move:
        for j in range(fat.shape[0]):
            if j != i and overlap(fat[i], fat[j]):
                out.add((min(i, j), max(i, j)))
    return out


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_pair_set_equals_brute_force(seed):
    rng = np.random.default_rng(seed)
    n = 300
    lo = rng.uniform(0, 30, (n, 2))
    ext = rng.uniform(0.1, 3.0, (n, 2))
    fat = np.concatenate([lo, lo + ext], axis=1)
    fat[5] = fat[6]  # an exact duplicate and touching edges (inclusive overlap)
    fat[7, 0] = fat[8, 2]
    move = rng.choice(n, size=60, replace=False).astype(np.int32)
    got = O.pairs(fat, move)
    assert {(min(a, b), max(a, b)) for a, b in got.tolist()} == _brute_pairs(fat, move.tolist())
    # canonical order: by move-buffer position of the first proxy, then ascending partner
    pos = {int(p): k for k, p in enumerate(move.tolist())}
    keys = [(pos[a], b) for a, b in got.tolist()]
    assert keys == sorted(keys)
    assert len(set(map(tuple, got.tolist()))) == len(got)


def _near_tie_state():
    import os
    from conftest import GOLDEN
    z = np.load(os.path.join(GOLDEN, "near_tie_mixed3975.npz"))
    st = L.BatchArrays(*[z[k] for k in L._FIELDS])
    st.world_i[:, L.WI_FLAGS] &= ~L.WF_NEW_FIXTURE
    return st


def _contact(st, fa, fb):
    k = np.nonzero((st.con_i[:, L.CI_FA] == fa) & (st.con_i[:, L.CI_FB] == fb))[0]
    assert len(k) == 1
    return st.con_i[k[0]]


def test_near_tie_fixture_flips_with_fp32_rounded_shapes():
    """The recorded state sits on the touching threshold of contact (229, 290): exact shapes
    give one touching point, the same shapes rounded to FP32 give none.  This is why the
    device keeps shape geometry in FP64 in both pipelines."""
    st = _near_tie_state()
    o = O.Oracle(st)
    o.step(DT, 8, 3)
    exact = _contact(o.state(), 229, 290)
    geo = [L.FD_RADIUS, L.FD_CENTROID_X, L.FD_CENTROID_Y, L.FD_CIRCLE_PX, L.FD_CIRCLE_PY] + \
        list(range(L.FD_EDGE_V0X, L.FD_EDGE_V3Y + 1)) + list(range(L.FD_VERTS, L.FD_NORMALS + 16))
    rounded = st.copy()
    rounded.fix_d[:, geo] = rounded.fix_d[:, geo].astype(np.float32).astype(np.float64)
    o2 = O.Oracle(rounded)
    o2.step(DT, 8, 3)
    flipped = _contact(o2.state(), 229, 290)
    assert exact[L.CI_PC] == 1 and exact[L.CI_FLAGS] & L.CF_TOUCHING
    assert flipped[L.CI_PC] == 0 and not flipped[L.CI_FLAGS] & L.CF_TOUCHING
