"""World building on the host equals the reference's, bit for bit.

The recorded initial states (tests/golden, written by pypybox2d itself) hold
mass data, centres of mass, polygon normals recomputed by the fixture copy,
the velocity shift of reset_mass_data and the initial fat AABBs.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN
from pypybox2d_amd import builder, layout as L, scenes

SCENES = {
    "pyramid_600.npz": lambda: scenes.pyramid(),
    "stacks_600.npz": lambda: scenes.stacks(seed=0),
    "mixed24_300.npz": lambda: scenes.mixed_world(np.random.default_rng(7), 24, half_width=8.0, height=12.0),
}


@pytest.mark.parametrize("run", sorted(SCENES))
def test_initial_state_matches_reference(run):
    z = np.load(os.path.join(GOLDEN, run))
    b = builder.scene_batch(SCENES[run]())
    for name in L._FIELDS:
        got, want = getattr(b, name), z["init_" + name]
        assert got.shape == want.shape, name
        assert np.array_equal(got, want), (name, np.argwhere(got != want)[:5].tolist())


def test_parallel_batch_builder_equals_serial():
    counts = scenes.mixed_batch_counts(6, seed=9, lo=8, hi=20)
    par = builder.batch_from_counts(counts, 9, procs=2, chunk=2)
    ser = builder.scenes_batch([scenes.batch_world(9, w, n) for w, n in enumerate(counts)])
    for name in L._FIELDS:
        assert np.array_equal(getattr(par, name), getattr(ser, name)), name
    assert par.n_worlds == 6
