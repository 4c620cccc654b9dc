"""The oracle's contact listener and custom contact filter against a recorded reference run.

Fixture listeners_filtered_200.npz (tests/golden/make_golden.py): the
filtered scene (category / mask / group bits and a kinematic platform) stepped
200 times by pypybox2d with every listener set on its contact manager and a
custom contact_filter (scenes.listener_filter_verdict: the default filter on
the bits and a rule on the fixture indices that is not symmetric, so a call
with swapped arguments would answer differently).  Its pre_solve disables
contacts by scenes.listener_disable_rule.  Fixture filter bits change three
times (Fixture.refilter): after step 60 a fixture stops colliding, after 110
it collides again, after 150 two touching fixtures join a negative group.
Every listener call is recorded in call order.

The oracle replays the run with the recorded contact-creation order, the
filter given as its verdict on every ordered fixture pair, and the pre_solve
verdicts applied between collide and the solve; it must reproduce every call
(kind, fixtures, flags, manifold, old manifold, impulses) bit for bit, in the
same order, and the recorded states.
"""
import os

import numpy as np

from conftest import GOLDEN
from oracle import oracle as O
from pypybox2d_amd import layout as L
from pypybox2d_amd.scenes import listener_disable_rule, listener_filter_verdict

RUN = os.path.join(GOLDEN, "listeners_filtered_200.npz")
# record fields the reference fixes (the world, phase and key only order the device's records)
CMP = [c for c in range(L.EV_NI) if c not in (L.EI_WORLD, L.EI_PHASE, L.EI_KEY, L.EI_PAD)]


def accept_matrix(fix_i):
    bits = [(int(r[L.FI_CATEGORY]), int(r[L.FI_MASK]), int(r[L.FI_GROUP])) for r in fix_i]
    n = len(bits)
    acc = np.zeros((n, n), np.uint8)
    for a in range(n):
        for b in range(n):
            acc[a, b] = listener_filter_verdict(a, bits[a], b, bits[b])
    return acc


def load(z, prefix):
    b = L.BatchArrays(*[z[prefix + n] for n in L._FIELDS])
    o = O.Oracle(b)
    o.set_events(L.EV_ALL)
    o.set_filter(0, accept_matrix(b.fix_i))
    return o


def replay(z):
    """Steps the oracle through the recorded run; yields (step, oracle, records of the step)."""
    hoff, hp = z["hint_off"], z["hint_pairs"]
    edit_steps = set(int(s) for s in z["edits"][:, 0])
    o = load(z, "init_")
    for s in range(1, int(z["meta_steps"]) + 1):
        if s - 1 in edit_steps:  # the filter bits changed after the last step: continue from that state
            o = load(z, "e%d_" % (s - 1))
        o.set_hint(0, hp[hoff[s - 1]:hoff[s]])
        assert o.step_begin(0) == 0
        ei, _ = o.events(0)
        for r in ei:
            if r[L.EI_KIND] == L.EV_PRE_SOLVE and listener_disable_rule(s, int(r[L.EI_FA]), int(r[L.EI_FB])):
                o.disable(0, int(r[L.EI_KEY]) // 4)
        assert o.step_end(0) == 0
        yield s, o, o.events(0)


def test_oracle_reproduces_every_listener_call_of_the_reference():
    z = np.load(RUN)
    off, gi, gd = z["ev_off"], z["ev_i"], z["ev_d"]
    cks = set(int(s) for s in z["checkpoints"])
    seen = np.zeros(4, int)
    disabled = 0
    for s, o, (ei, ed) in replay(z):
        want_i, want_d = gi[off[s - 1]:off[s]], gd[off[s - 1]:off[s]]
        assert ei.shape[0] == want_i.shape[0], (s, ei.shape[0], want_i.shape[0])
        assert np.array_equal(ei[:, CMP], want_i[:, CMP]), (s, np.argwhere(ei[:, CMP] != want_i[:, CMP])[:5].tolist())
        assert np.array_equal(ed, want_d), (s, np.argwhere(ed != want_d)[:5].tolist())
        np.add.at(seen, ei[:, L.EI_KIND], 1)
        disabled += sum(1 for r in ei if r[L.EI_KIND] == L.EV_PRE_SOLVE and
                        listener_disable_rule(s, int(r[L.EI_FA]), int(r[L.EI_FB])))
        if s in cks:
            st = o.state()
            for name in ("body_d", "fix_d", "con_d", "con_i", "move"):
                got, want = getattr(st, name), z["s%d_%s" % (s, name)]
                assert np.array_equal(got, want), (s, name)
    assert np.all(seen > 20) and disabled > 10, (seen, disabled)


def test_the_run_exercises_the_filter_and_the_destroying_end_contact():
    """The custom filter rejects pairs the default filter admits (so the verdicts matter), the
    refilters destroy touching contacts (end_contact from ContactManager.destroy: an end call
    while the contact still reads touching), and some post_solve calls carry two impulses."""
    z = np.load(RUN)
    fi = z["init_fix_i"]
    acc = accept_matrix(fi)
    bits = [(int(r[L.FI_CATEGORY]), int(r[L.FI_MASK]), int(r[L.FI_GROUP])) for r in fi]

    def default(a, b):
        (ca, ma, ga), (cb, mb, gb) = bits[a], bits[b]
        return ga > 0 if (ga == gb and ga != 0) else ((ma & cb) != 0 and (ca & mb) != 0)
    n = len(bits)
    assert any(default(a, b) and not acc[a, b] for a in range(n) for b in range(n))
    assert any(acc[a, b] != acc[b, a] for a in range(n) for b in range(n))
    gi = z["ev_i"]
    ends = gi[gi[:, L.EI_KIND] == L.EV_END]
    assert np.any(ends[:, L.EI_FLAGS] & L.CF_TOUCHING) and np.any(~ends[:, L.EI_FLAGS] & L.CF_TOUCHING)
    posts = gi[gi[:, L.EI_KIND] == L.EV_POST_SOLVE]
    assert np.any(posts[:, L.EI_OLD_PC] == 2) and np.any(posts[:, L.EI_OLD_PC] == 1)
    assert len(z["edits"]) == 4
