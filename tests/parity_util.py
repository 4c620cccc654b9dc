"""Comparison of two world states in the batch layout (GPU result vs oracle).

Relative error of a quantity q is |q_gpu - q_ref| / max(|q_ref|, floor_q) with
the floors: position 1 m, angle 1 rad, linear velocity |g| dt, angular
velocity |g| dt / 1 m, accumulated impulses m_min |g| dt (m_min = the
smaller non-zero mass of the contact's two bodies).
"""
import numpy as np

from pypybox2d_amd import layout as L

G_DT = 10.0 / 60.0


def body_errors(gpu: L.BatchArrays, ref: L.BatchArrays):
    g, r = gpu.body_d, ref.body_d
    out = {}

    def rel(cols, floor):
        a, b = g[:, cols], r[:, cols]
        return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor))) if a.size else 0.0
    out["position"] = rel([L.BD_C_X, L.BD_C_Y], 1.0)
    out["angle"] = rel([L.BD_A], 1.0)
    out["velocity"] = rel([L.BD_V_X, L.BD_V_Y], G_DT)
    out["angular_velocity"] = rel([L.BD_W], G_DT)
    out["xf"] = rel([L.BD_XF_PX, L.BD_XF_PY], 1.0)
    out["rotation"] = rel([L.BD_XF_C, L.BD_XF_S], 1.0)
    return out


def contact_keys(st: L.BatchArrays):
    """Per world: list of (fa, fb) in contact order."""
    keys = []
    for w in range(st.n_worlds):
        c0, c1 = st.con_off[w], st.con_off[w + 1]
        keys.append([tuple(x) for x in st.con_i[c0:c1, :2].tolist()])
    return keys


def contact_errors(gpu: L.BatchArrays, ref: L.BatchArrays):
    """Impulse errors over contacts present in both (matched by world and fixture pair)."""
    bmass = ref.body_d[:, L.BD_MASS]
    worst = {"normal_impulse": 0.0, "tangent_impulse": 0.0}
    flips = 0
    matched = 0
    for w in range(ref.n_worlds):
        b0, f0 = ref.body_off[w], ref.fix_off[w]
        rc = {tuple(ref.con_i[k, :2]): k for k in range(ref.con_off[w], ref.con_off[w + 1])}
        for k in range(gpu.con_off[w], gpu.con_off[w + 1]):
            key = tuple(gpu.con_i[k, :2])
            if key not in rc:
                flips += 1
                continue
            j = rc[key]
            gi, ri = gpu.con_i[k], ref.con_i[j]
            if gi[L.CI_PC] != ri[L.CI_PC] or gi[L.CI_MTYPE] != ri[L.CI_MTYPE] or \
                    (gi[L.CI_FLAGS] & 3) != (ri[L.CI_FLAGS] & 3):
                flips += 1
                continue
            matched += 1
            ba = b0 + ref.fix_i[f0 + ri[L.CI_FA], L.FI_BODY]
            bb = b0 + ref.fix_i[f0 + ri[L.CI_FB], L.FI_BODY]
            ms = [m for m in (bmass[ba], bmass[bb]) if m > 0]
            floor = (min(ms) if ms else 1.0) * G_DT
            pc = int(ri[L.CI_PC])
            for p, (ni, ti) in enumerate(((L.CD_P0_NI, L.CD_P0_TI), (L.CD_P1_NI, L.CD_P1_TI))[:pc]):
                for name, col in (("normal_impulse", ni), ("tangent_impulse", ti)):
                    e = abs(gpu.con_d[k, col] - ref.con_d[j, col]) / max(abs(ref.con_d[j, col]), floor)
                    worst[name] = max(worst[name], float(e))
    worst["discrete_flips"] = flips
    worst["matched_contacts"] = matched
    return worst


def same_contact_lists(a: L.BatchArrays, b: L.BatchArrays) -> bool:
    return contact_keys(a) == contact_keys(b)


QUANTITIES = ("position", "angle", "velocity", "angular_velocity")
IMPULSES = ("normal_impulse", "tangent_impulse")


def assert_one_step(got: L.BatchArrays, ref: L.BatchArrays, tol: float, what=""):
    """The one-step bar: equal contact lists, no contact whose discrete state (point count,
    manifold type, touching/enabled) differs, and every body quantity and accumulated impulse
    within `tol` relative (floors as in the module docstring).  Returns both error dicts."""
    assert same_contact_lists(got, ref), ("contact lists differ", what)
    be, ce = body_errors(got, ref), contact_errors(got, ref)
    print("  one step", what, {k: "%.2e" % be[k] for k in QUANTITIES},
          {k: "%.2e" % ce[k] for k in IMPULSES}, "matched", ce["matched_contacts"], flush=True)
    assert ce["discrete_flips"] == 0, (what, ce)
    for q in QUANTITIES:
        assert be[q] <= tol, (what, q, be[q])
    for q in IMPULSES:
        assert ce[q] <= tol, (what, q, ce[q])
    return be, ce


def golden_state(z, init, s, dt=1.0 / 60):
    """The full batch at checkpoint s of a recorded reference run (a one-world npz written by
    tests/golden/make_golden.py), joints included when the run has them."""
    b = init.copy()
    b.body_d, b.body_i = z["s%d_body_d" % s].copy(), z["s%d_body_i" % s].copy()
    b.fix_d = z["s%d_fix_d" % s].copy()
    b.con_d, b.con_i = z["s%d_con_d" % s].copy(), z["s%d_con_i" % s].copy()
    b.move = z["s%d_move" % s].copy()
    b.con_off = np.array([0, b.con_d.shape[0]], np.int32)
    b.move_off = np.array([0, b.move.shape[0]], np.int32)
    if "s%d_joint_d" % s in z.files:
        b.joint_d, b.joint_i = z["s%d_joint_d" % s].copy(), z["s%d_joint_i" % s].copy()
    b.world_d[0, L.WD_INV_DT0] = 1.0 / dt
    b.world_i[0, L.WI_FLAGS] &= ~L.WF_NEW_FIXTURE
    return b
