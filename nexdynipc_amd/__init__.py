"""MI355X (gfx950) implementation of NexDynIPC's contact stage of the implicit-Euler
Newton step, behind a C ABI (include/ndipc.h).

    from nexdynipc_amd import scenes, Engine, IPCSolver, World, RigidBody, load_scene

Engine is the batch engine over the C ABI; IPCSolver / World / RigidBody / FixedJoint /
HingeJoint mirror the reference's solver interface on top of it; load_scene reads the
reference's JSON scenes (include/ndipc_io.h).
"""
from . import scenes  # noqa: F401

_LAZY = {
    "Engine": "_lib", "NdipcError": "_lib", "make_params": "_lib",
    "IPCSolver": "solver", "World": "solver", "RigidBody": "solver", "FixedJoint": "solver",
    "HingeJoint": "solver", "load_scene": "scene_io", "Exporter": "scene_io",
}


def __getattr__(name):
    mod = _LAZY.get(name)
    if mod is None:
        raise AttributeError(name)
    import importlib

    return getattr(importlib.import_module("." + mod, __name__), name)
