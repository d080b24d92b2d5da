"""Geometric delays of the ATA phasor module, computed on the host.

``ata_delays`` calls ``blade_ata_delays`` of the C-ABI, the same code the
``Phasor::ATA`` module runs per block (reference
src/modules/phasor/ata.cc:105-173): beam-minus-boresight delays in seconds,
``[B, A]``.
"""

from __future__ import annotations

import ctypes as ct

import numpy as np

from . import _lib


def ata_delays(antenna_ecef, reference_lla, boresight_radec, beam_radec, julian_date,
               dut1=0.0, reference_antenna=0):
    """Delays [B, A] (s) of every beam and antenna relative to the reference
    antenna, minus the boresight's.

    antenna_ecef: [A, 3] metres; reference_lla: (lon, lat, alt) in rad, rad, m;
    boresight_radec: (ra, dec) rad, J2000; beam_radec: [B, 2] rad, J2000;
    julian_date: UTC; dut1: UT1 - UTC in seconds.
    """
    ecef = np.ascontiguousarray(antenna_ecef, dtype=np.float64).reshape(-1, 3)
    beams = np.ascontiguousarray(beam_radec, dtype=np.float64).reshape(-1, 2)
    lla = np.ascontiguousarray(reference_lla, dtype=np.float64).reshape(3)
    bore = np.ascontiguousarray(boresight_radec, dtype=np.float64).reshape(2)
    A, B = ecef.shape[0], beams.shape[0]
    out = np.zeros((B, A), dtype=np.float64)
    ptr = lambda a: a.ctypes.data_as(ct.c_void_p)  # noqa: E731
    _lib.call("blade_ata_delays", ptr(ecef), A, int(reference_antenna), ptr(lla), ptr(bore),
              ptr(beams), B, float(julian_date), float(dut1), ptr(out))
    return out
