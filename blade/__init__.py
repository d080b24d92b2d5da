"""``import blade as bl``: the reference package's import name
(python/blade/__init__.py:1-12) for this library, so scripts written for
the reference run unchanged.

Everything is ``blade_amd``'s: the constants, ``array_tensor`` and the other
tensors, ``copy``, ``module``, ``pipeline`` and ``runner``.  The extension
submodules are reachable under both names (``blade._beamformer_impl`` is
``blade_amd._beamformer_impl``), as are ``blade._internal`` (the constants,
as the reference's ``blade._internal``) and ``blade._runner_impl``.
"""

import sys as _sys

import blade_amd as _amd
from blade_amd import *  # noqa: F401,F403
from blade_amd._api import __all__  # noqa: F401

for _name, _mod in list(_sys.modules.items()):
    if _name.startswith("blade_amd._") and _name.endswith("_impl"):
        _sys.modules["blade." + _name[len("blade_amd."):]] = _mod
_sys.modules["blade._internal"] = _amd

del _name, _mod
