"""blade_amd: the ATA beamforming path of Blade on MI355X (gfx950).

Python mirror of the reference package ``blade`` (python/blade/__init__.py):
constants (``bl.cf32``, ``bl.cuda``, ``bl.ata``, ``bl.beamformer`` ...),
tensors (``bl.array_tensor`` ...), ``bl.module``, ``@bl.runner``,
``bl.copy``.  Everything below runs on the C++ API of ``libblade_amd.so``
through the ``_blade_ext`` pybind11 extension; there is no CPU fallback.
"""

from __future__ import annotations

import importlib
import inspect as _inspect
import sys as _sys
from types import SimpleNamespace

# One HIP runtime per process: PyTorch's libraries ask for the unversioned
# ``libamdhip64.so`` (its bundled copy) while libblade_amd.so asks for the
# SONAME ``libamdhip64.so.7``.  Loading torch first makes that SONAME resolve
# to torch's copy; loading this library first would bring in a second runtime
# that torch then duplicates.
try:
    import torch as _torch  # noqa: F401
except ImportError:
    _torch = None

try:
    from . import _blade_ext as _ext
except ImportError as _e:
    # Only the build recipe (blade_amd.build) and the raw C-ABI helpers
    # (_lib, ops) are usable without the extension; every API name raises.
    _import_error = ImportError(
        "blade_amd._blade_ext is missing or failed to load; build it with "
        "`python -m blade_amd.build` (the HIP library is required): " + str(_e))

    def __getattr__(name):
        # Submodules stay importable (``from blade_amd import build`` looks
        # the name up here before it imports the submodule).
        if name in ("build", "_lib", "ops", "autotune", "calibration", "delays"):
            return importlib.import_module(f"{__name__}.{name}")
        raise _import_error

    _ext = None

if _ext is not None:
    from ._api import *  # noqa: F401,F403
    from ._api import _Constant, _fetch, _sanitize_duet  # noqa: F401
