"""MI355X (gfx950) implementation of NexDynIPC's contact stage of the implicit-Euler
Newton step, behind a C ABI (include/ndipc.h).

    from nexdynipc_amd import scenes, Engine, IPCSolver
"""
from . import scenes  # noqa: F401


def __getattr__(name):
    if name in ("Engine", "NdipcError", "make_params"):
        from . import _lib
        return getattr(_lib, name)
    if name == "IPCSolver":
        from .solver import IPCSolver
        return IPCSolver
    raise AttributeError(name)
