"""The solver seam, b2g_solve_islands: the island search and Island.solve (world.py:494-580,
island.py:65-179, contact.py:375-864, the joints) alone, against the oracle's orc_solve_islands,
from recorded reference states.

No narrowphase runs (the manifolds are the recorded ones) and no pair search, so a solver
difference shows here apart from any narrowphase or contact-order difference.  Bars: FP64 every
body quantity, accumulated impulse and joint-record value within 1e-12 (relative, floors as in
parity_util); FP32 within 1e-5, the one-step bar, impulses included; the solve leaves the contact
list and the fat AABBs as they were.
"""
import os

import numpy as np
import pytest

import parity_util as PU
from conftest import GOLDEN
from oracle import oracle as O
from pypybox2d_amd import _native as N
from pypybox2d_amd import layout as L

pytestmark = pytest.mark.gpu

DT = 1.0 / 60
RUNS = [("pyramid_600.npz", (10, 100, 600)), ("stacks_600.npz", (30, 300)), ("mixed24_300.npz", (40, 300)),
        ("filtered_kinematic_300.npz", (60, 250)), ("joints_bridge_300.npz", (20, 200)),
        ("joints_machines_300.npz", (20, 150, 300))]


def seam(state, prec):
    e = N.Engine(prec)
    try:
        e.upload(state)
        e.solve_islands(DT, 8, 3)
        return e.download()
    finally:
        e.close()


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-5)])
@pytest.mark.parametrize("run,checkpoints", RUNS)
def test_solve_islands_matches_the_oracle(run, checkpoints, prec, tol):
    z = np.load(os.path.join(GOLDEN, run))
    init = L.load_batch(z, "init_")
    for s in checkpoints:
        state = PU.golden_state(z, init, s)
        o = O.Oracle(state)
        assert o.solve_islands(DT, 8, 3) == 0
        ref, got = o.state(), seam(state, prec)
        PU.assert_one_step(got, ref, tol, "%s step %d %s" % (run, s, prec))
        assert np.array_equal(got.con_i[:, :2], state.con_i[:, :2])  # no contact created or destroyed
        assert np.array_equal(got.fix_d, state.fix_d)                 # no fixture synchronised
        if state.n_joints:
            je = float(np.max(np.abs(got.joint_d - ref.joint_d) / np.maximum(np.abs(ref.joint_d), 1.0)))
            assert je <= (1e-12 if prec == "fp64" else 1e-5), (s, je)
            assert np.array_equal(got.joint_i, ref.joint_i), s


def test_solve_islands_rejects_misuse():
    z = np.load(os.path.join(GOLDEN, "pyramid_600.npz"))
    e = N.Engine("fp64")
    e.upload(L.load_batch(z, "init_"))
    for bad in ((0.0, 8, 3), (DT, -1, 3)):
        with pytest.raises(N.B2GError) as x:
            e.solve_islands(*bad)
        assert x.value.code == N.B2G_E_INVALID
    e.step_collide(DT, 8, 3)
    with pytest.raises(N.B2GError):
        e.solve_islands(DT, 8, 3)
    e.step_solve()
    e.solve_islands(DT, 8, 3)
    e.close()
