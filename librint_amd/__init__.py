"""librint_amd: MI355X (gfx950) HIP implementation of librint's two-electron
hot path -- Rys-quadrature ERIs fused with the J/K contraction of the RHF
Fock build -- behind the reference's C ABI (include/librint.h).

Submodules: mol (libcint arrays and workload builders), scf (the reference's
Python API), context (persistent GPU context), build (hipcc build).
"""
from . import mol  # noqa: F401

__all__ = ["mol", "scf", "context"]


def __getattr__(name):
    if name in ("scf", "context", "utils"):
        import importlib
        return importlib.import_module("." + name, __name__)
    raise AttributeError(name)
