"""Writes the seeded golden fixtures of the beamforming path.

Each fixture holds the inputs and the complex128 oracle's outputs, stored in
the dtypes the HIP path produces (complex64 / float32), at shapes small enough
to keep every file well under 1 MB:

* ``mode_b_small.npz``     -- ATA Mode-B beamformer at the small CPU case
  (A20 T512 P2 B8, incoherent beam on), channels 0..3 of the 192-channel band:
  CI8 N(0, 16^2) voltages, tau ~ U(+-50 ns), f_k = 1 GHz + k * 0.5 MHz, CF32
  phasors (reference ``tests/modules/beamformer/ata.py:56-63``).
* ``chain_rate4.npz``      -- the production Mode-B chain at A28 B16 on 2
  channels x 512 samples: beamform -> FFT-4 channelizer -> XY->LR polarizer.
* ``channelizer.npz``      -- rate 4 and rate == T channelizer cases
  (reference ``tests/modules/channelizer/advanced.py:49-65``).
* ``polarizer_detector.npz`` -- polarizer and 1-/4-pol detector outputs
  (``tests/modules/polarizer/polarizer.py:47-49``,
  ``tests/modules/detector/single_pol.py:58-67``, ``cross_pol.py:58-71``).
* ``phasor.npz``           -- phasors from delays and per-pol calibration with
  the ATA observation parameters of ``tests/modules/phasor/advanced.py``
  (f_obs 6500.125 MHz, channel bandwidth 0.5 MHz, total bandwidth 1.024 GHz,
  frequency start index 352).
* ``bitexact.npz``         -- Cast, Gather and Permutation inputs/outputs.

Run from the repository root: ``python tests/golden/make_golden.py``.
The oracle itself is pinned independently by ``tests/test_oracle.py``.
"""

from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle as O  # noqa: E402


def _crand(rng, *shape):
    return (rng.normal(size=shape) + 1j * rng.normal(size=shape)).astype(np.complex64)


def mode_b_small():
    rng = np.random.default_rng(0)
    A, F, T, P, B = 20, 4, 512, 2, 8
    x = O.ci8_gaussian((A, F, T, P), rng=rng)
    tau = rng.uniform(-50e-9, 50e-9, size=(B, A))
    cal = np.ones((A, F, 1, P), dtype=np.complex128)
    w = O.phasor_ata(tau, cal, 1.0e9, 0.5e6, 0.0, 0).astype(np.complex64)
    y = O.beamformer_ata(x, w, enable_incoherent_beam=True)
    return dict(x=x, tau=tau, w=w, y=y.astype(np.complex64))


def chain_rate4():
    rng = np.random.default_rng(1)
    A, F, T, P, B = 28, 2, 512, 2, 16
    x = O.ci8_gaussian((A, F, T, P), rng=rng)
    tau = rng.uniform(-50e-9, 50e-9, size=(B, A))
    cal = np.ones((A, F, 1, P), dtype=np.complex128)
    w = O.phasor_ata(tau, cal, 1.0e9 + 100 * 0.5e6, 0.5e6, 0.0, 0).astype(np.complex64)
    y = O.polarizer(O.channelizer(O.beamformer_ata(x, w), 4))
    return dict(x=x, w=w, y=y.astype(np.complex64))


def channelizer():
    rng = np.random.default_rng(2)
    x4 = _crand(rng, 2, 3, 64, 2)
    xT = _crand(rng, 2, 3, 256, 2)
    return dict(x4=x4, y4=O.channelizer(x4, 4).astype(np.complex64),
                xT=xT, yT=O.channelizer(xT, 256).astype(np.complex64))


def polarizer_detector():
    rng = np.random.default_rng(3)
    x = _crand(rng, 2, 3, 32, 2)
    return dict(x=x, lr=O.polarizer(x).astype(np.complex64),
                d1=O.detector(x, 4, 1).astype(np.float32),
                d4=O.detector(x, 2, 4).astype(np.float32))


def phasor():
    rng = np.random.default_rng(4)
    B, A, F, P = 2, 20, 8, 2
    tau = rng.uniform(-200e-9, 200e-9, size=(B, A))
    cal = (rng.normal(size=(A, F, 1, P)) + 1j * rng.normal(size=(A, F, 1, P)))
    w = O.phasor_ata(tau, cal, 6500.125e6, 0.5e6, 1.024e9, 352)
    return dict(tau=tau, cal=cal, w=w, f_obs=6500.125e6, chan_bw=0.5e6, total_bw=1.024e9,
                fstart=352)


def bitexact():
    rng = np.random.default_rng(5)
    ci8 = rng.integers(-128, 128, size=(3, 4, 16, 2, 2), dtype=np.int8)
    blocks = [_crand(rng, 2, 3, 8, 2) for _ in range(3)]
    perm_in = _crand(rng, 2, 3, 16, 2)
    return dict(ci8=ci8, ci8_cf32=O.cast(ci8, "ci8", "cf32"),
                g0=blocks[0], g1=blocks[1], g2=blocks[2], gather_t=O.gather(blocks, 2),
                perm_in=perm_in, perm_0132=O.permutation(perm_in, (0, 1, 3, 2)),
                perm_0213=O.permutation(perm_in, (0, 2, 1, 3)))


FIXTURES = {
    "mode_b_small": mode_b_small,
    "chain_rate4": chain_rate4,
    "channelizer": channelizer,
    "polarizer_detector": polarizer_detector,
    "phasor": phasor,
    "bitexact": bitexact,
}


def main():
    for name, fn in FIXTURES.items():
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **fn())
        print(f"{path}: {os.path.getsize(path)} bytes")


if __name__ == "__main__":
    main()
