"""The Runner's asynchronous mode with distinct blocks (reference
src/runner.cc:17-68, benchmarks/bundles/ata/mode-b/mode_b.hh:127-154).

enqueue() puts successive blocks on alternating streams, two in flight, while
every module keeps one output buffer.  Each test feeds a different block
(and, for Mode B, a different Julian date) per enqueue and checks every
output against the oracle of exactly its blocks, so a cycle that reads or
writes a buffer of the neighbouring cycle fails it.
"""

import math
import time

import numpy as np
import pytest

import blade_amd as bl
import oracle as O
from blade_amd.delays import ata_delays
from test_gpu_api import ATA_POS, crand, fill, modeb_config

pytestmark = pytest.mark.gpu

TOL = 1e-5


def drive(p, n_blocks, n_outputs, transfer_in, transfer_out):
    """Enqueue n_blocks through the Runner, dequeue until n_outputs are out."""
    state = {"in": 0, "out": 0, "done": 0}

    def input_callback():
        transfer_in(state["in"])
        state["in"] += 1
        return bl.result.success

    def output_callback():
        transfer_out(state["out"])
        state["out"] += 1
        return bl.result.success

    def callback(input_id, output_id, did_output):
        if did_output:
            state["done"] += 1
        return bl.result.success

    # dequeue() polls without blocking once fewer blocks than streams are in
    # flight (as the reference's does), so the loop is bounded by time, not
    # by a count of polls.
    deadline = time.monotonic() + 120.0
    while time.monotonic() < deadline:
        if state["in"] == n_blocks and state["done"] == n_outputs:
            break
        if state["in"] < n_blocks:
            p.enqueue(input_callback, output_callback, state["in"], state["out"])
        p.dequeue(callback)
    assert state["in"] == n_blocks and state["done"] == n_outputs, state


@bl.runner
class SteppedDetector:
    def __init__(self, in_shape, out_shape, intg):
        self.input.buf = bl.array_tensor(in_shape, dtype=bl.cf32)
        self.output.buf = bl.array_tensor(out_shape, dtype=bl.f32)
        self.module.detector = bl.module(bl.detector, {"integration_size": intg,
                                                       "number_of_output_polarizations": 1},
                                         self.input.buf, ot=bl.f32)

    def transfer_in(self, buf):
        self.copy(self.input.buf, buf)

    def transfer_out(self, buf):
        self.copy(self.output.buf, self.module.detector.get_output())
        self.copy(buf, self.output.buf)


def test_stepped_detector_distinct_blocks():
    # Integration over two blocks: the detector adds the second block of each
    # pair into its output; eight different blocks give four outputs.
    shape = (8, 64, 2048, 2)
    out_shape = (8, 64, 1, 1)
    n = 8
    rng = np.random.default_rng(101)
    blocks = [crand(rng, shape) for _ in range(n)]
    host_in = [bl.array_tensor(shape, dtype=bl.cf32, device=bl.cpu) for _ in range(n)]
    for h, x in zip(host_in, blocks):
        fill(h, x)
    host_out = [bl.array_tensor(out_shape, dtype=bl.f32, device=bl.cpu) for _ in range(n // 2)]
    p = SteppedDetector(shape, out_shape, 2 * shape[2])
    drive(p, n, n // 2, lambda i: p.transfer_in(host_in[i]), lambda k: p.transfer_out(host_out[k]))
    for k in range(n // 2):
        ref = O.detector_stepped(blocks[2 * k:2 * k + 2], 2 * shape[2], 1)
        assert O.norm_rel_err(host_out[k].as_numpy(), ref) <= TOL, k


def test_mode_bh_distinct_blocks_and_dates():
    # tests/bundles/ata/mode_bh.py's composition (Mode B beams gathered 4x
    # along time into one Mode-H block), with a new voltage block and a
    # Julian date 10 minutes later on every enqueue.
    A, F, T = 20, 192, 512
    b_in, b_out = (A, F, T, 2), (2, F, T, 2)
    h_in, h_out = (2, F, 4 * T, 2), (2, F * 4 * T, 1, 1)
    n = 8
    rng = np.random.default_rng(202)
    cal_np = rng.random((A, F, 1, 2)) + 1j * rng.random((A, F, 1, 2))
    cal = bl.array_tensor((A, F, 1, 2), dtype=bl.cf64, device=bl.cpu)
    fill(cal, cal_np)
    config_b = modeb_config(b_in, b_out, cal, 1, False, False)
    config_b["beamformer_incoherent_beam"] = False
    config_h = {"input_shape": h_in, "output_shape": h_out, "polarizer_convert_to_circular": True,
                "detector_integration_size": 1, "detector_number_of_output_polarizations": 1}

    @bl.runner
    class Pipeline:
        def __init__(self):
            self.input.dut = bl.tensor(1, dtype=bl.f64, device=bl.cpu)
            self.input.date = bl.tensor(1, dtype=bl.f64, device=bl.cpu)
            self.input.buffer = bl.array_tensor(b_in, dtype=bl.cf32)
            self.output.buffer = bl.array_tensor(h_out, dtype=bl.f32)
            inp = (self.input.dut, self.input.date, self.input.buffer)
            self.module.mode_b = bl.module(bl.modeb, config_b, inp, telescope=bl.ata)
            self.module.gather = bl.module(bl.gather, {"axis": 2, "multiplier": 4},
                                           self.module.mode_b.get_output(), ot=bl.cf32)
            self.module.mode_h = bl.module(bl.modeh, config_h, self.module.gather.get_output(),
                                           ot=bl.f32)

        def transfer_in(self, dut, date, buffer):
            self.copy(self.input.dut, dut)
            self.copy(self.input.date, date)
            self.copy(self.input.buffer, buffer)

        def transfer_out(self, buffer):
            self.copy(self.output.buffer, self.module.mode_h.get_output())
            self.copy(buffer, self.output.buffer)

    date0 = (1649366473.0 / 86400) + 2440587.5
    dates = [date0 + i * 600.0 / 86400.0 for i in range(n)]
    xs = [crand(rng, b_in) for _ in range(n)]
    host = []
    for i in range(n):
        dut = bl.tensor(1, dtype=bl.f64, device=bl.cpu)
        date = bl.tensor(1, dtype=bl.f64, device=bl.cpu)
        buf = bl.array_tensor(b_in, dtype=bl.cf32, device=bl.cpu)
        dut.as_numpy()[0] = 0.0
        date.as_numpy()[0] = dates[i]
        fill(buf, xs[i])
        host.append((dut, date, buf))
    host_out = [bl.array_tensor(h_out, dtype=bl.f32, device=bl.cpu) for _ in range(n // 4)]
    p = Pipeline()
    drive(p, n, n // 4, lambda i: p.transfer_in(*host[i]), lambda k: p.transfer_out(host_out[k]))

    lla = (math.radians(-121.470733), math.radians(40.815987), 1020.86)
    beams = [(0.63722, 1.07552424), (0.65063, 1.08426835)]
    bore = (0.64169, 1.079896295)
    for k in range(n // 4):
        parts = []
        for i in range(4 * k, 4 * k + 4):
            d = ata_delays(np.array(ATA_POS[:A]), lla, bore, beams, dates[i])
            w = O.phasor_ata(d, cal_np, 6500.125e6, 0.5e6, 1.024e9, 352)
            parts.append(O.mode_b(xs[i], w))                      # [2, F, T, 2]
        ref = O.mode_h(np.concatenate(parts, axis=2), convert_to_circular=True)
        assert ref.shape == h_out
        assert O.norm_rel_err(host_out[k].as_numpy(), ref) <= TOL, k
    # the dates matter: the first output with the first block's phasors for
    # all four blocks is far off
    d0 = ata_delays(np.array(ATA_POS[:A]), lla, bore, beams, dates[0])
    w0 = O.phasor_ata(d0, cal_np, 6500.125e6, 0.5e6, 1.024e9, 352)
    stale = O.mode_h(np.concatenate([O.mode_b(xs[i], w0) for i in range(4)], axis=2),
                     convert_to_circular=True)
    assert O.norm_rel_err(host_out[0].as_numpy(), stale) > 1e-3
