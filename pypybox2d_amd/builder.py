"""Scene specs (pypybox2d_amd.scenes) -> drop-in worlds and batch records."""
from __future__ import annotations

from typing import List

from . import layout as L
from .fixture import Fixture
from .shapes import Circle, Edge, Polygon
from .scenes import fixture_filter
from .world import World


def make_shape(spec):
    kind = spec[0]
    if kind == "box":
        return Polygon(box=(spec[1], spec[2]))
    if kind == "circle":
        return Circle(radius=spec[1])
    if kind == "polygon":
        return Polygon(vertices=spec[1])
    if kind == "edge":
        return Edge(spec[1], spec[2])
    raise ValueError("unknown shape %r" % (kind,))


def build_world(scene, precision: str = "fp64", device: int = 0) -> World:
    """Builds the scene through the public API, in the spec's body order."""
    world = World(gravity=scene["gravity"], do_sleep=scene["do_sleep"], warm_starting=scene["warm_starting"],
                  continuous_physics=scene.get("continuous", False), precision=precision, device=device)
    for bs in scene["bodies"]:
        fixtures = [Fixture(make_shape(f["shape"]), density=f["density"], friction=f["friction"],
                            restitution=f["restitution"], category_bits=fixture_filter(f)[0],
                            mask_bits=fixture_filter(f)[1], group_index=fixture_filter(f)[2])
                    for f in bs["fixtures"]]
        kw = dict(position=bs["position"], angle=bs["angle"], linear_velocity=bs["linear_velocity"],
                  angular_velocity=bs["angular_velocity"], fixtures=fixtures, bullet=bs.get("bullet", False))
        if bs["type"] == "static":
            world.create_static_body(**kw)
        elif bs["type"] == "kinematic":
            world.create_kinematic_body(**kw)
        else:
            world.create_dynamic_body(**kw)
    made = []
    for js in scene.get("joints", ()):
        kw = {k: v for k, v in js.items() if k not in ("type", "bodies", "joints")}
        create = getattr(world, "create_%s_joint" % js["type"])
        if js["type"] == "gear":
            made.append(create(made[js["joints"][0]], made[js["joints"][1]], **kw))
        else:
            made.append(create(*[world.bodies[i] for i in js["bodies"]], **kw))
    return world


def scene_batch(scene) -> L.BatchArrays:
    """Initial state of one scene in the batch layout (host only, no device)."""
    return build_world(scene)._batch()


def scenes_batch(scenes: List[dict]) -> L.BatchArrays:
    return L.concat([scene_batch(s) for s in scenes])


def _build_range(args):
    seed, counts, w0 = args
    from . import scenes
    return scenes_batch([scenes.batch_world(seed, w0 + k, n) for k, n in enumerate(counts)])


def batch_from_counts(counts, seed: int, procs: int = 0, chunk: int = 64) -> L.BatchArrays:
    """Builds the worlds of a seeded batch (scenes.batch_world) in worker processes.

    Call it before the process touches the GPU: the workers are forked."""
    import multiprocessing as mp
    import os
    counts = [int(c) for c in counts]
    jobs = [(seed, counts[i:i + chunk], i) for i in range(0, len(counts), chunk)]
    procs = procs or min(16, os.cpu_count() or 1)
    if procs <= 1 or len(jobs) == 1:
        parts = [_build_range(j) for j in jobs]
    else:
        with mp.get_context("fork").Pool(procs) as pool:
            parts = pool.map(_build_range, jobs)
    return L.concat(parts)
