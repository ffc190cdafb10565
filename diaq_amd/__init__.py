"""DiaQ GPU path on MI355X (gfx950): diagonal-format sparse linear algebra with hand-written HIP kernels.

``diaq_amd.diaq_gpu`` is the drop-in for the reference's ``diaq_gpu`` module (importing this package also makes
``import diaq_gpu`` work); ``diaq_amd.capi`` is the ctypes view of the C ABI (include/diaq_hip.h). Both need the
native libraries built in-tree (``python diaq_amd/build.py``): there is no CPU fallback, a missing library is an
ImportError.
"""
import os
import sys

_HERE = os.path.dirname(os.path.abspath(__file__))

# One HIP runtime per process: torch ships its own libamdhip64 (same soname as /opt/rocm's). Loading torch first makes
# this library bind to the runtime torch already loaded; loading this library first would pull in the system runtime
# and leave torch with a second, GPU-less one.
try:
    import torch  # noqa: F401
except ImportError:  # torch is optional: without it the system HIP runtime is used
    pass

try:
    from . import diaq_gpu  # noqa: F401  (links libdiaq_hip.so next to it)
except ImportError as e:  # pragma: no cover - depends on the build
    raise ImportError(f"diaq_amd: native module not built or not loadable ({e}); run python diaq_amd/build.py") from e

sys.modules.setdefault("diaq_gpu", diaq_gpu)

LIBRARY = os.path.join(_HERE, "libdiaq_hip.so")
