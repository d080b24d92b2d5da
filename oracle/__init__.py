"""CPU oracle for the ATA beamforming path — TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker.  The product
path (``blade_amd``) never imports it and fails loudly when its HIP library is
missing.

Every function restates, in our own numpy code, the algorithm of the
reference file:line cited in its docstring.  Each one is pinned by
independent checks in ``tests/test_oracle.py`` (brute-force loops, closed
forms, invariants); see the "Pinned by" line of each docstring.
"""

from .blade_oracle import *  # noqa: F401,F403
