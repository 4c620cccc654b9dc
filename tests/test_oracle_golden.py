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

RUNS = ["pyramid_600.npz", "stacks_600.npz", "mixed24_300.npz"]


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
