"""ctypes wrapper of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
(as the checker) and by ``bench.py``'s cpu_baseline leg.  The product package
``pypybox2d_amd`` never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from pypybox2d_amd import layout as L

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "all"], cwd=_HERE, check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        l = ctypes.CDLL(_LIB_PATH)
        vp, i32, f64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_double
        l.orc_load.restype = vp
        l.orc_load.argtypes = [ctypes.POINTER(L.CBatch)]
        l.orc_free.argtypes = [vp]
        l.orc_step.argtypes = [vp, f64, ctypes.c_int, ctypes.c_int]
        l.orc_step_world.argtypes = [vp, ctypes.c_int, f64, ctypes.c_int, ctypes.c_int]
        l.orc_counts.argtypes = [vp, vp]
        l.orc_save.argtypes = [vp, ctypes.POINTER(L.CBatchOut)]
        l.orc_set_hint.argtypes = [vp, ctypes.c_int, vp, ctypes.c_int]
        l.orc_error.argtypes = [vp]
        l.orc_evaluate.argtypes = [vp] * 8
        l.orc_pairs.argtypes = [ctypes.c_int, vp, ctypes.c_int, vp, vp, ctypes.c_int]
        l.orc_compute_aabb.argtypes = [vp, vp, vp, vp]
        l.orc_shape_ray_cast.argtypes = [vp, vp, vp, vp, f64, vp]
        l.orc_query_aabb.argtypes = [vp, ctypes.c_int, vp, vp, ctypes.c_int]
        l.orc_ray_cast.argtypes = [vp, ctypes.c_int, vp, vp, vp, ctypes.c_int]
        l.orc_set_events.argtypes = [vp, ctypes.c_int]
        l.orc_events.argtypes = [vp, ctypes.c_int, vp, vp, ctypes.c_int]
        l.orc_set_filter.argtypes = [vp, ctypes.c_int, vp, ctypes.c_int]
        l.orc_new_contacts.argtypes = [vp, ctypes.c_int, vp, ctypes.c_int]
        l.orc_step_begin.argtypes = [vp, ctypes.c_int, f64, ctypes.c_int, ctypes.c_int]
        l.orc_disable.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        l.orc_step_end.argtypes = [vp, ctypes.c_int]
        l.orc_toi_stats.argtypes = [vp, ctypes.c_int, vp]
        l.orc_solve_islands.argtypes = [vp, f64, ctypes.c_int, ctypes.c_int]
        _ = i32
        _lib = l
    return _lib


ORC_E_UNSUPPORTED, ORC_E_HINT, ORC_E_ALLOC, ORC_E_TOI = 1, 2, 4, 8


class Oracle:
    """A batch of worlds stepped by the oracle."""

    def __init__(self, batch: L.BatchArrays):
        self._batch = batch.normalized()
        self._struct = self._batch.as_struct()
        self.h = lib().orc_load(ctypes.byref(self._struct))
        if not self.h:
            raise ValueError("oracle rejected the batch")

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            lib().orc_free(h)
            self.h = None

    def step(self, dt=1.0 / 60, vel_iters=8, pos_iters=3) -> int:
        return lib().orc_step(self.h, dt, vel_iters, pos_iters)

    def solve_islands(self, dt=1.0 / 60, vel_iters=8, pos_iters=3) -> int:
        """The island search and Island.solve alone on the current state (the solver seam)."""
        return lib().orc_solve_islands(self.h, dt, vel_iters, pos_iters)

    def step_world(self, w: int, dt=1.0 / 60, vel_iters=8, pos_iters=3) -> int:
        return lib().orc_step_world(self.h, w, dt, vel_iters, pos_iters)

    def set_hint(self, w: int, pairs) -> None:
        p = np.ascontiguousarray(np.asarray(pairs, np.int32).reshape(-1, 2))
        rc = lib().orc_set_hint(self.h, w, p.ctypes.data if p.size else None, int(p.shape[0]))
        if rc:
            raise RuntimeError("orc_set_hint failed")

    def step_begin(self, w: int, dt=1.0 / 60, vel_iters=8, pos_iters=3) -> int:
        """World w's step up to and with collide (the listener's begin/end/pre_solve calls)."""
        rc = lib().orc_step_begin(self.h, w, dt, vel_iters, pos_iters)
        if rc < 0:
            raise ValueError("orc_step_begin: bad world or a step already begun")
        return rc

    def disable(self, w: int, slot: int) -> None:
        """Contact.enabled = False for the contact at index `slot` during this step's collide."""
        if lib().orc_disable(self.h, w, int(slot)) < 0:
            raise ValueError("orc_disable: no such contact")

    def step_end(self, w: int) -> int:
        rc = lib().orc_step_end(self.h, w)
        if rc < 0:
            raise ValueError("orc_step_end: no step begun")
        return rc

    def toi_stats(self, w: int = 0):
        """(time_of_impact calls, TOI events) of world w's last step (continuous physics)."""
        out = np.zeros(2, np.int64)
        if lib().orc_toi_stats(self.h, w, out.ctypes.data):
            raise ValueError("bad world")
        return int(out[0]), int(out[1])

    def set_events(self, mask: int) -> None:
        """Records the listener calls of the kinds in mask (1 << layout.EV_*), for every world."""
        lib().orc_set_events(self.h, int(mask))

    def events(self, w: int):
        """The last step's listener records of world w in call order: (int32 rows of EV_NI, float64 rows
        of EV_ND)."""
        cap = 256
        while True:
            ei = np.zeros((cap, L.EV_NI), np.int32)
            ed = np.zeros((cap, L.EV_ND))
            n = lib().orc_events(self.h, w, ei.ctypes.data, ed.ctypes.data, cap)
            if n == -1:
                raise ValueError("bad world")
            if n >= 0:
                return ei[:n].copy(), ed[:n].copy()
            cap *= 4

    def set_filter(self, w: int, accept) -> None:
        """A custom contact_filter for world w as its verdicts accept[a, b] = filter(fixture a, fixture b)
        (world-local fixtures); None restores World.default_contact_filter."""
        if accept is None:
            rc = lib().orc_set_filter(self.h, w, None, 0)
        else:
            a = np.ascontiguousarray(accept, np.uint8)
            rc = lib().orc_set_filter(self.h, w, a.ctypes.data, int(a.shape[0]))
        if rc:
            raise ValueError("orc_set_filter failed")

    def new_contacts(self, w: int) -> np.ndarray:
        """Rows (fixture A, fixture B, flags) of the contacts world w's last pair search created."""
        cap = 256
        while True:
            out = np.zeros((cap, 3), np.int32)
            n = lib().orc_new_contacts(self.h, w, out.ctypes.data, cap)
            if n == -1:
                raise ValueError("bad world")
            if n >= 0:
                return out[:n].copy()
            cap *= 4

    def query_aabb(self, w: int, box) -> np.ndarray:
        """World.query_aabb by brute force: fixtures of world w whose fat AABB overlaps box."""
        b = np.ascontiguousarray(box, np.float64)
        cap = 64
        while True:
            out = np.zeros(cap, np.int32)
            n = lib().orc_query_aabb(self.h, w, b.ctypes.data, out.ctypes.data, cap)
            if n < 0:
                raise ValueError("bad world")
            if n <= cap:
                return out[:n].copy()
            cap = n

    def ray_cast(self, w: int, seg):
        """World.ray_cast, every hit unclipped, in fixture order: (fixtures, rows of fraction, point x, y,
        normal x, y)."""
        sg = np.ascontiguousarray(seg, np.float64)
        cap = 64
        while True:
            of = np.zeros(cap, np.int32)
            od = np.zeros((cap, 5))
            n = lib().orc_ray_cast(self.h, w, sg.ctypes.data, of.ctypes.data, od.ctypes.data, cap)
            if n < 0:
                raise ValueError("p1 == p2 or bad world")
            if n <= cap:
                return of[:n].copy(), od[:n].copy()
            cap = n

    def error(self) -> int:
        return lib().orc_error(self.h)

    def counts(self):
        c = np.zeros(4, np.int32)
        lib().orc_counts(self.h, c.ctypes.data)
        return c

    def state(self) -> L.BatchArrays:
        c = self.counts()
        out = L.OutBuffers(self._batch, con_cap=int(c[2]) + 16, move_cap=int(c[3]) + 16)
        rc = lib().orc_save(self.h, ctypes.byref(out.struct))
        if rc:
            raise RuntimeError("orc_save failed: %d" % rc)
        return out.to_batch()


def evaluate(fa_d, fa_i, fb_d, fb_i, xfa, xfb):
    """Manifold of one contact (fixtures oriented A, B) -> (con_d row, con_i row)."""
    cd = np.zeros(L.CON_ND)
    ci = np.zeros(L.CON_NI, np.int32)
    args = [np.ascontiguousarray(fa_d, np.float64), np.ascontiguousarray(fa_i, np.int32),
            np.ascontiguousarray(fb_d, np.float64), np.ascontiguousarray(fb_i, np.int32),
            np.ascontiguousarray(xfa, np.float64), np.ascontiguousarray(xfb, np.float64)]
    rc = lib().orc_evaluate(*[a.ctypes.data for a in args], cd.ctypes.data, ci.ctypes.data)
    if rc:
        raise ValueError("no collider for this shape pair")
    return cd, ci


def pairs(fat, move):
    """Canonical new-pair list of the broad-phase for fat AABBs (n x 4) and a move buffer."""
    fat = np.ascontiguousarray(fat, np.float64).reshape(-1, 4)
    move = np.ascontiguousarray(move, np.int32)
    cap = max(16, fat.shape[0] * max(1, move.shape[0]))
    out = np.zeros((cap, 2), np.int32)
    n = lib().orc_pairs(fat.shape[0], fat.ctypes.data, move.shape[0],
                        move.ctypes.data if move.size else None, out.ctypes.data, cap)
    if n < 0:
        raise RuntimeError("orc_pairs failed: %d" % n)
    return out[:n].copy()


def compute_aabb(fd, fi, xf):
    out = np.zeros(4)
    a = [np.ascontiguousarray(fd, np.float64), np.ascontiguousarray(fi, np.int32),
         np.ascontiguousarray(xf, np.float64)]
    lib().orc_compute_aabb(a[0].ctypes.data, a[1].ctypes.data, a[2].ctypes.data, out.ctypes.data)
    return out


def shape_ray_cast(fd, fi, xf, seg, max_fraction=1.0):
    """Shape.ray_cast of one fixture's shape at transform xf = (px, py, c, s): (hit, fraction, nx, ny)."""
    out = np.zeros(3)
    a = [np.ascontiguousarray(fd, np.float64), np.ascontiguousarray(fi, np.int32),
         np.ascontiguousarray(xf, np.float64), np.ascontiguousarray(seg, np.float64)]
    hit = lib().orc_shape_ray_cast(a[0].ctypes.data, a[1].ctypes.data, a[2].ctypes.data, a[3].ctypes.data,
                                   float(max_fraction), out.ctypes.data)
    return bool(hit), float(out[0]), float(out[1]), float(out[2])
