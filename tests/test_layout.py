"""include/b2g_layout.h and pypybox2d_amd/layout.py describe the same records."""
import os
import re

from conftest import REPO
from pypybox2d_amd import layout as L


def _header_enums():
    text = open(os.path.join(REPO, "include", "b2g_layout.h")).read()
    values = {"B2G_MAX_POLY_VERTS": 8}
    for body in re.findall(r"enum\s*\{(.*?)\}", text, re.S):
        body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
        cur = -1
        for item in [t.strip() for t in body.split(",") if t.strip()]:
            if "=" in item:
                name, expr = [s.strip() for s in item.split("=", 1)]
                cur = eval(expr, {}, dict(values))
            else:
                name, cur = item, cur + 1
            values[name] = cur
    for name, val in re.findall(r"#define\s+(B2G_\w+)\s+(\d+)u?", text):
        values[name] = int(val)
    return values


def test_layout_constants_match_header():
    h = _header_enums()
    checked = 0
    for name, val in h.items():
        py = name[len("B2G_"):]
        if hasattr(L, py):
            assert getattr(L, py) == val, name
            checked += 1
    assert checked > 80
    assert h["B2G_FD_NORMALS"] + 2 * h["B2G_MAX_POLY_VERTS"] <= h["B2G_FIX_ND"]
    assert h["B2G_BD_SLEEP_TIME"] < h["B2G_BODY_ND"]


def test_id_pack_roundtrip():
    for t in [(0, 0, 0, 0), (7, 3, 1, 0), (1, 0, 0, 1), (255, 255, 1, 1)]:
        assert L.unpack_id(L.pack_id(*t)) == t
