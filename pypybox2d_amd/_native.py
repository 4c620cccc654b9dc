"""ctypes binding of libb2g.so (the gfx950 HIP library, include/b2g.h).

The library is loaded from this package directory.  There is no fallback: if
it is missing or the GPU is absent, the calls raise.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

from . import layout as L

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libb2g.so")

B2G_OK = 0
B2G_E_INVALID, B2G_E_CAPACITY, B2G_E_HIP, B2G_E_NAN, B2G_E_UNSUPPORTED, B2G_E_NODEVICE = -1, -2, -3, -4, -5, -6
B2G_FP32, B2G_FP64 = 0, 1

EXPORTS = ("b2g_version", "b2g_create", "b2g_destroy", "b2g_upload", "b2g_step", "b2g_step_n", "b2g_sync",
           "b2g_counts", "b2g_download", "b2g_last_error", "b2g_set_param", "b2g_get_param", "b2g_profile_step",
           "b2g_island_times", "b2g_narrowphase", "b2g_pairs", "b2g_stats", "b2g_query_aabb", "b2g_ray_cast",
           "b2g_pairs_noalloc", "b2g_events", "b2g_pair_search", "b2g_new_contacts", "b2g_drop_contacts",
           "b2g_step_collide", "b2g_step_solve", "b2g_disable_contacts", "b2g_solve_islands")

# keys of b2g_set_param / b2g_get_param: the solver's launch shape, island timing, and the
# resident solver wavefronts per compute unit and the search launches per step (both read only)
PARAMS = {"team_lds": 1, "heavy_class": 2, "solve_grid": 3, "team_full_class": 4, "large_worlds": 5,
          "island_times": 6, "solve_waves_per_cu": 7, "team_max": 8, "search_runs": 9, "pairs_wide": 10,
          "islands_wide": 11, "collide_block": 12, "tree_pairs": 13, "events": 14}

# the stages b2g_profile_step times, in order
STAGES = ("collide", "islands", "solve", "sync", "pairs")

_ERRNAMES = {B2G_E_INVALID: "invalid argument", B2G_E_CAPACITY: "capacity exceeded", B2G_E_HIP: "HIP error",
             B2G_E_NAN: "non-finite state", B2G_E_UNSUPPORTED: "unsupported feature",
             B2G_E_NODEVICE: "no GPU device"}


class B2GError(RuntimeError):
    def __init__(self, code: int, msg: str = ""):
        self.code = code
        super().__init__("b2g error %d (%s)%s" % (code, _ERRNAMES.get(code, "?"), (": " + msg) if msg else ""))


class CDesc(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int32), ("device", ctypes.c_int32), ("contact_factor", ctypes.c_double),
                ("contact_extra", ctypes.c_int32), ("reserved", ctypes.c_int32)]


_lib = None


def load(path: Optional[str] = None):
    """Load libb2g.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise RuntimeError("libb2g.so is not built; run `python -m pypybox2d_amd.build`")
    lib = ctypes.CDLL(path)
    vp, i32, f64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_double
    lib.b2g_version.restype = ctypes.c_int
    lib.b2g_create.argtypes = [ctypes.POINTER(CDesc), ctypes.POINTER(vp)]
    lib.b2g_destroy.argtypes = [vp]
    lib.b2g_upload.argtypes = [vp, ctypes.POINTER(L.CBatch)]
    lib.b2g_step.argtypes = [vp, f64, ctypes.c_int, ctypes.c_int]
    lib.b2g_step_n.argtypes = [vp, ctypes.c_int, f64, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.b2g_sync.argtypes = [vp]
    lib.b2g_counts.argtypes = [vp, vp]
    lib.b2g_download.argtypes = [vp, ctypes.POINTER(L.CBatchOut)]
    lib.b2g_last_error.argtypes = [vp, ctypes.c_char_p, ctypes.c_size_t]
    lib.b2g_profile_step.argtypes = [vp, ctypes.c_int, f64, ctypes.c_int, ctypes.c_int, vp]
    lib.b2g_narrowphase.argtypes = [ctypes.c_int, ctypes.c_int] + [vp] * 8
    lib.b2g_pairs.argtypes = [ctypes.c_int, vp, ctypes.c_int, vp, vp, ctypes.c_int, vp]
    lib.b2g_set_param.argtypes = [vp, ctypes.c_int32, ctypes.c_int32]
    lib.b2g_get_param.argtypes = [vp, ctypes.c_int32, vp]
    lib.b2g_island_times.argtypes = [vp, vp, ctypes.c_int32, vp]
    lib.b2g_stats.argtypes = [vp, vp, ctypes.c_int32, ctypes.c_int32]
    lib.b2g_query_aabb.argtypes = [vp, ctypes.c_int, vp, vp, vp, ctypes.c_int, vp]
    lib.b2g_ray_cast.argtypes = [vp, ctypes.c_int, vp, vp, ctypes.c_int32, vp, vp, ctypes.c_int, vp]
    lib.b2g_events.argtypes = [vp, vp, vp, ctypes.c_int32, vp]
    lib.b2g_pair_search.argtypes = [vp]
    lib.b2g_new_contacts.argtypes = [vp, vp, ctypes.c_int32, vp]
    lib.b2g_drop_contacts.argtypes = [vp, ctypes.c_int32, vp, vp]
    lib.b2g_step_collide.argtypes = [vp, f64, ctypes.c_int, ctypes.c_int]
    lib.b2g_step_solve.argtypes = [vp]
    lib.b2g_disable_contacts.argtypes = [vp, ctypes.c_int32, vp, vp]
    lib.b2g_solve_islands.argtypes = [vp, f64, ctypes.c_int, ctypes.c_int]
    _ = i32
    _lib = lib
    return lib


class Engine:
    """A batch of worlds resident on one GPU (a b2g_ctx)."""

    def __init__(self, precision: str = "fp32", device: int = 0, contact_factor: float = 8.0,
                 contact_extra: int = 64):
        self.lib = load()
        self.precision = precision
        code = {"fp32": B2G_FP32, "fp64": B2G_FP64}[precision]
        desc = CDesc(code, device, float(contact_factor), int(contact_extra), 0)
        h = ctypes.c_void_p()
        rc = self.lib.b2g_create(ctypes.byref(desc), ctypes.byref(h))
        if rc:
            raise B2GError(rc, "b2g_create")
        self.h = h
        self._like: Optional[L.BatchArrays] = None

    def _check(self, rc: int, what: str) -> None:
        if rc:
            buf = ctypes.create_string_buffer(512)
            self.lib.b2g_last_error(self.h, buf, 512)
            raise B2GError(rc, "%s: %s" % (what, buf.value.decode(errors="replace")))

    def close(self) -> None:
        h = getattr(self, "h", None)
        if h is not None and h.value:
            self.lib.b2g_destroy(h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, batch: L.BatchArrays) -> None:
        b = batch.normalized()
        s = b.as_struct()
        self._check(self.lib.b2g_upload(self.h, ctypes.byref(s)), "upload")
        self._like = b

    def step(self, dt: float = 1.0 / 60, vel_iters: int = 8, pos_iters: int = 3) -> None:
        self._check(self.lib.b2g_step(self.h, dt, vel_iters, pos_iters), "step")

    def step_n(self, n: int, dt: float = 1.0 / 60, vel_iters: int = 8, pos_iters: int = 3,
               graph: bool = True) -> None:
        self._check(self.lib.b2g_step_n(self.h, n, dt, vel_iters, pos_iters, 1 if graph else 0), "step_n")

    def sync(self) -> None:
        self._check(self.lib.b2g_sync(self.h), "sync")

    def counts(self) -> np.ndarray:
        c = np.zeros(4, np.int32)
        self._check(self.lib.b2g_counts(self.h, c.ctypes.data), "counts")
        return c

    def download(self) -> L.BatchArrays:
        c = self.counts()
        out = L.OutBuffers(self._like, con_cap=int(c[2]) + 16, move_cap=int(c[1]) + 16)
        self._check(self.lib.b2g_download(self.h, ctypes.byref(out.struct)), "download")
        return out.to_batch()

    def set_param(self, name: str, value: int) -> None:
        """Solver launch shape: team_lds, heavy_class, team_full_class, team_max or solve_grid (results do not
        depend on them); island_times; large_worlds."""
        self._check(self.lib.b2g_set_param(self.h, PARAMS[name], int(value)), "set_param")

    def get_param(self, name: str) -> int:
        v = np.zeros(1, np.int32)
        self._check(self.lib.b2g_get_param(self.h, PARAMS[name], v.ctypes.data), "get_param")
        return int(v[0])

    def profile(self, n: int, dt: float = 1.0 / 60, vel_iters: int = 8, pos_iters: int = 3) -> np.ndarray:
        ms = np.zeros(5, np.float32)
        self._check(self.lib.b2g_profile_step(self.h, n, dt, vel_iters, pos_iters, ms.ctypes.data), "profile")
        return ms

    def stats(self, reset: bool = False) -> dict:
        """Work counters summed over the steps since the last reset (b2g_stats)."""
        v = np.zeros(9, np.int64)
        self._check(self.lib.b2g_stats(self.h, v.ctypes.data, 9, 1 if reset else 0), "stats")
        return {"steps": int(v[0]), "live_contacts": int(v[1]), "evaluated_contacts": int(v[2]),
                "new_contacts": int(v[3]), "moved_proxies": int(v[4]), "island_bodies": int(v[5]),
                "island_contacts": int(v[6]), "toi_calls": int(v[7]), "toi_events": int(v[8])}

    def query_aabb(self, worlds, boxes) -> np.ndarray:
        """World.query_aabb for many boxes: rows (query, world-local fixture) of every fixture whose fat
        AABB overlaps box k = (lower x, lower y, upper x, upper y) of world worlds[k], by query then
        fixture."""
        w = np.ascontiguousarray(worlds, np.int32).reshape(-1)
        b = np.ascontiguousarray(boxes, np.float64).reshape(-1, 4)
        if w.shape[0] != b.shape[0]:
            raise ValueError("one world index per box")
        cap = max(64, 8 * w.shape[0])
        while True:
            out = np.zeros((cap, 2), np.int32)
            n = np.zeros(1, np.int32)
            rc = self.lib.b2g_query_aabb(self.h, w.shape[0], w.ctypes.data, b.ctypes.data, out.ctypes.data, cap,
                                         n.ctypes.data)
            if rc == B2G_E_CAPACITY and n[0] > cap:
                cap = int(n[0])
                continue
            self._check(rc, "query_aabb")
            return out[:int(n[0])].copy()

    def ray_cast(self, worlds, segments, closest: bool = True):
        """World.ray_cast for many segments (p1 x, p1 y, p2 x, p2 y) of worlds[k].  Returns (rows, data):
        rows (ray, world-local fixture; -1 for no hit), data (fraction, point x, y, normal x, y); one
        row per ray (its closest hit) or every hit sorted by ray, fraction, fixture."""
        w = np.ascontiguousarray(worlds, np.int32).reshape(-1)
        sg = np.ascontiguousarray(segments, np.float64).reshape(-1, 4)
        if w.shape[0] != sg.shape[0]:
            raise ValueError("one world index per segment")
        cap = max(w.shape[0], 64 if closest else 8 * w.shape[0])
        while True:
            oi = np.zeros((cap, 2), np.int32)
            od = np.zeros((cap, 5))
            n = np.zeros(1, np.int32)
            rc = self.lib.b2g_ray_cast(self.h, w.shape[0], w.ctypes.data, sg.ctypes.data, 1 if closest else 0,
                                       oi.ctypes.data, od.ctypes.data, cap, n.ctypes.data)
            if rc == B2G_E_CAPACITY and n[0] > cap:
                cap = int(n[0])
                continue
            self._check(rc, "ray_cast")
            m = int(n[0])
            return oi[:m].copy(), od[:m].copy()

    def events(self):
        """The contact listener's records of the last step (b2g_events; kinds enabled with
        set_param("events", mask)), in the reference's call order: (int32 rows of layout.EV_NI,
        float64 rows of layout.EV_ND)."""
        n = np.zeros(1, np.int32)
        cap = 256
        while True:
            ei = np.zeros((cap, L.EV_NI), np.int32)
            ed = np.zeros((cap, L.EV_ND))
            rc = self.lib.b2g_events(self.h, ei.ctypes.data, ed.ctypes.data, cap, n.ctypes.data)
            if rc == B2G_E_CAPACITY and n[0] > cap:
                cap = int(n[0])
                continue
            self._check(rc, "events")
            m = int(n[0])
            return ei[:m].copy(), ed[:m].copy()

    def step_collide(self, dt: float = 1.0 / 60, vel_iters: int = 8, pos_iters: int = 3) -> None:
        """The first part of a step, up to and with the narrow phase (b2g_step_collide); its listener
        records are then readable (events) and contacts can be disabled for the step."""
        self._check(self.lib.b2g_step_collide(self.h, dt, vel_iters, pos_iters), "step_collide")

    def disable_contacts(self, worlds, slots) -> None:
        """contact.enabled = False for this step: slot slots[k] (a phase-0 record's key // 4) of world
        worlds[k], between step_collide and step_solve (b2g_disable_contacts)."""
        w = np.ascontiguousarray(worlds, np.int32).reshape(-1)
        j = np.ascontiguousarray(slots, np.int32).reshape(-1)
        if w.shape[0] != j.shape[0]:
            raise ValueError("one world index per slot")
        if w.shape[0] == 0:
            return
        self._check(self.lib.b2g_disable_contacts(self.h, w.shape[0], w.ctypes.data, j.ctypes.data), "disable_contacts")

    def step_solve(self) -> None:
        """The rest of the step begun by step_collide (b2g_step_solve)."""
        self._check(self.lib.b2g_step_solve(self.h), "step_solve")

    def solve_islands(self, dt: float, vel_iters: int, pos_iters: int) -> None:
        """The island search and Island.solve alone on the current state (b2g_solve_islands): no
        collide, no fixture synchronisation, no pair search (a parity seam for the solver)."""
        self._check(self.lib.b2g_solve_islands(self.h, float(dt), int(vel_iters), int(pos_iters)), "solve_islands")

    def pair_search(self) -> None:
        """Runs the pending new-fixture pair search ahead of the next step (b2g_pair_search)."""
        self._check(self.lib.b2g_pair_search(self.h), "pair_search")

    def new_contacts(self) -> np.ndarray:
        """Rows (world, fixture A, fixture B, flags) of the contacts each world's last pair search created
        (b2g_new_contacts); flags: layout.CF_SWAPPED, layout.NC_DROPPED."""
        n = np.zeros(1, np.int32)
        cap = 64
        while True:
            out = np.zeros((cap, 4), np.int32)
            rc = self.lib.b2g_new_contacts(self.h, out.ctypes.data, cap, n.ctypes.data)
            if rc == B2G_E_CAPACITY and n[0] > cap:
                cap = int(n[0])
                continue
            self._check(rc, "new_contacts")
            return out[:int(n[0])].copy()

    def drop_contacts(self, worlds, index) -> None:
        """Removes new contacts index[k] of world worlds[k] (indices among that world's rows of
        new_contacts) before the next step (b2g_drop_contacts)."""
        w = np.ascontiguousarray(worlds, np.int32).reshape(-1)
        j = np.ascontiguousarray(index, np.int32).reshape(-1)
        if w.shape[0] != j.shape[0]:
            raise ValueError("one world index per contact")
        if w.shape[0] == 0:
            return
        self._check(self.lib.b2g_drop_contacts(self.h, w.shape[0], w.ctypes.data, j.ctypes.data), "drop_contacts")

    def island_times(self) -> np.ndarray:
        """Islands of the last step (island timing on): rows of start_ns, end_ns, contacts,
        bodies, lanes, sweep period, velocity_end_ns, position iterations."""
        n = np.zeros(1, np.int32)
        self._check(self.lib.b2g_island_times(self.h, None, 0, n.ctypes.data), "island_times")
        raw = np.zeros((max(int(n[0]), 1), 6), np.int64)
        self._check(self.lib.b2g_island_times(self.h, raw.ctypes.data, int(n[0]), n.ctypes.data), "island_times")
        raw = raw[:int(n[0])]
        out = np.zeros((raw.shape[0], 8), np.int64)
        out[:, 0], out[:, 1] = raw[:, 0], raw[:, 1]
        out[:, 2], out[:, 3] = raw[:, 2] & 0xFFFFFFFF, raw[:, 2] >> 32
        out[:, 4], out[:, 5] = raw[:, 3] & 0xFFFFFFFF, raw[:, 3] >> 32
        out[:, 6], out[:, 7] = raw[:, 4], raw[:, 5]
        return out


def narrowphase(precision: str, fa_d, fa_i, fb_d, fb_i, xfa, xfb):
    """Parity seam: manifolds of n oriented fixture pairs (rows of the records)."""
    lib = load()
    arrs = [np.ascontiguousarray(fa_d, np.float64).reshape(-1, L.FIX_ND),
            np.ascontiguousarray(fa_i, np.int32).reshape(-1, L.FIX_NI),
            np.ascontiguousarray(fb_d, np.float64).reshape(-1, L.FIX_ND),
            np.ascontiguousarray(fb_i, np.int32).reshape(-1, L.FIX_NI),
            np.ascontiguousarray(xfa, np.float64).reshape(-1, 4),
            np.ascontiguousarray(xfb, np.float64).reshape(-1, 4)]
    n = arrs[0].shape[0]
    cd = np.zeros((n, L.CON_ND))
    ci = np.zeros((n, L.CON_NI), np.int32)
    rc = lib.b2g_narrowphase({"fp32": B2G_FP32, "fp64": B2G_FP64}[precision], n,
                             *[a.ctypes.data for a in arrs], cd.ctypes.data, ci.ctypes.data)
    if rc:
        raise B2GError(rc, "b2g_narrowphase")
    return cd, ci


def pairs(fat, move):
    """Parity seam: canonical broad-phase pair list for fat AABBs and a move buffer."""
    lib = load()
    fat = np.ascontiguousarray(fat, np.float64).reshape(-1, 4)
    move = np.ascontiguousarray(move, np.int32)
    cap = max(16, fat.shape[0] * max(1, move.shape[0]))
    out = np.zeros((cap, 2), np.int32)
    cnt = np.zeros(1, np.int32)
    rc = lib.b2g_pairs(fat.shape[0], fat.ctypes.data, move.shape[0], move.ctypes.data if move.size else None,
                       out.ctypes.data, cap, cnt.ctypes.data)
    if rc:
        raise B2GError(rc, "b2g_pairs")
    return out[:int(cnt[0])].copy()
