"""The FP64 oracle against the reference's own recorded behaviour.

The fixtures under tests/golden/ were written by tests/golden/make_golden.py
from pypybox2d (pure-Python flavour).  The oracle replays each run with the
reference's recorded contact-creation order and must match every recorded
state bit for bit: bodies (sweep, transform, velocities, sleep time, flags),
fat AABBs, the ordered contact list with its manifolds and impulses, and the
broad-phase move buffer.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN
from oracle import oracle as O
from pypybox2d_amd import layout as L

RUNS = ["pyramid_600.npz", "stacks_600.npz", "mixed24_300.npz", "filtered_kinematic_300.npz"]


def load_init(z):
    return L.BatchArrays(*[z["init_" + n] for n in L._FIELDS])


def replay(path, check_every=True):
    z = np.load(path)
    orc = O.Oracle(load_init(z))
    hoff, hp = z["hint_off"], z["hint_pairs"]
    cks = set(int(s) for s in z["checkpoints"])
    for s in range(1, int(z["meta_steps"]) + 1):
        orc.set_hint(0, hp[hoff[s - 1]:hoff[s]])
        assert orc.step() == 0, "oracle error word at step %d" % s
        if s in cks:
            st = orc.state()
            for name in ("body_d", "fix_d", "con_d", "con_i", "move"):
                got, want = getattr(st, name), z["s%d_%s" % (s, name)]
                assert got.shape == want.shape, (s, name, got.shape, want.shape)
                assert np.array_equal(got, want), (s, name, np.argwhere(got != want)[:5].tolist())
            assert np.array_equal(st.body_i[:, :2], z["s%d_body_i" % s][:, :2]), s
    return orc


@pytest.mark.parametrize("run", RUNS)
def test_oracle_replays_reference_bit_exact(run):
    replay(os.path.join(GOLDEN, run))


def test_oracle_narrowphase_cases_bit_exact():
    z = np.load(os.path.join(GOLDEN, "narrowphase.npz"))
    n = z["fa_d"].shape[0]
    assert n >= 500
    touching = 0
    for k in range(n):
        cd, ci = O.evaluate(z["fa_d"][k], z["fa_i"][k], z["fb_d"][k], z["fb_i"][k], z["xfa"][k], z["xfb"][k])
        assert np.array_equal(ci, z["con_i"][k]), (k, str(z["kinds"][k]))
        assert np.array_equal(cd, z["con_d"][k]), (k, str(z["kinds"][k]))
        touching += int(ci[L.CI_PC] > 0)
    assert touching > n // 2  # the cases exercise the contact branches, not just early-outs


def test_pyramid_run_reaches_sleep_quirk():
    """The 600-step pyramid run crosses the island sleep test (island.py:161-179):
    some checkpoint has a dynamic body put to sleep by the stale-variable rule."""
    z = np.load(os.path.join(GOLDEN, "pyramid_600.npz"))
    slept = False
    for s in z["checkpoints"]:
        bi = z["s%d_body_i" % s]
        dyn = bi[:, L.BI_TYPE] == L.DYNAMIC
        slept |= bool(np.any((bi[dyn, L.BI_FLAGS] & L.BF_AWAKE) == 0))
    bd = z["s600_body_d"]
    assert np.all(np.isfinite(bd))
    # sleeping or not, the pyramid stands: top box above 9 m
    assert bd[:, L.BD_C_Y].max() > 9.0
    assert slept or np.any(bd[1:, L.BD_SLEEP_TIME] > 0)


def _default_filter(fa, fb):
    """world.py:958-970 (the reference's default contact filter) on (category, mask, group)."""
    if fa[2] == fb[2] and fa[2] != 0:
        return fa[2] > 0
    return (fa[1] & fb[0]) != 0 and (fa[0] & fb[1]) != 0


def test_filtered_run_exercises_every_filter_branch_and_a_kinematic_body():
    """The recorded run with collision filtering and a kinematic platform: at its checkpoints
    there are fat-AABB-overlapping pairs that the mask bits reject, that a negative group rejects
    and that a positive group admits against the masks, and none of the rejected pairs is a
    contact, while every admitted overlapping pair with a dynamic body is; the kinematic body
    moves and touches dynamic bodies."""
    z = np.load(os.path.join(GOLDEN, "filtered_kinematic_300.npz"))
    fi = z["init_fix_i"]
    bi0 = z["init_body_i"]
    filt = [(int(r[L.FI_CATEGORY]), int(r[L.FI_MASK]), int(r[L.FI_GROUP])) for r in fi]
    kin = np.nonzero(bi0[:, L.BI_TYPE] == L.KINEMATIC)[0]
    assert len(kin) == 1
    seen = {"mask_reject": 0, "group_reject": 0, "group_admit": 0, "kinematic_touch": 0}
    for s in z["checkpoints"]:
        fat = z["s%d_fix_d" % s][:, [L.FD_FAT_LX, L.FD_FAT_LY, L.FD_FAT_UX, L.FD_FAT_UY]]
        ci = z["s%d_con_i" % s]
        cons = set((int(a), int(b)) for a, b in ci[:, :2]) | set((int(b), int(a)) for a, b in ci[:, :2])
        for a in range(len(fi)):
            for b in range(a + 1, len(fi)):
                ba, bb = fi[a, L.FI_BODY], fi[b, L.FI_BODY]
                if ba == bb or L.DYNAMIC not in (bi0[ba, L.BI_TYPE], bi0[bb, L.BI_TYPE]):
                    continue
                if fat[b, 0] - fat[a, 2] > 0 or fat[b, 1] - fat[a, 3] > 0 or fat[a, 0] - fat[b, 2] > 0 or \
                        fat[a, 1] - fat[b, 3] > 0:
                    continue
                ok = _default_filter(filt[a], filt[b])
                assert ((a, b) in cons) == ok, (int(s), a, b, filt[a], filt[b])
                fa, fb = filt[a], filt[b]
                if fa[2] == fb[2] and fa[2] != 0:
                    seen["group_admit" if fa[2] > 0 else "group_reject"] += int(((fa[1] & fb[0]) == 0) or fa[2] < 0)
                elif not ok:
                    seen["mask_reject"] += 1
        k_touch = [k for k in range(ci.shape[0]) if ci[k, L.CI_FLAGS] & L.CF_TOUCHING and
                   kin[0] in (fi[ci[k, L.CI_FA], L.FI_BODY], fi[ci[k, L.CI_FB], L.FI_BODY])]
        seen["kinematic_touch"] += len(k_touch)
    print(seen)
    assert all(v > 0 for v in seen.values()), seen
    bd = z["s300_body_d"]
    assert abs(bd[kin[0], L.BD_C_X] - 5.0) < 1e-9  # moved 300 steps at 1 m/s, untouched by contacts
