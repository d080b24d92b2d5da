"""Batched uneven blocks in one persistent launch, hipGraph capture, autotuner.

The batched launch runs the same per-tile code as the single-block launch
(the same tile widths and product forms with automatic tuning), so each
block's beams equal a single ``blade_beamformer_ata`` call bit for bit.
Forced tile widths change the product form and agree to rounding (normwise
1e-6).  A few channels of every block are also
checked against the complex128 oracle (normwise relative error <= 1e-5).
Graph replays are checked one by one against eager results, with new inputs
copied in between.
"""

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-5


def host(t):
    import torch
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def block_shapes(n, seed, Fs=(64, 128, 256, 1024), Ts=(1024, 2048, 4096, 8192)):
    """Seeded draw of uneven blocks: A in {14,20,27,28}, F, T powers of two, B in 4..64."""
    rng = np.random.default_rng(seed)
    return [(int(rng.choice([14, 20, 27, 28])), int(rng.choice(Fs)), int(rng.choice(Ts)),
             int(rng.integers(4, 65))) for _ in range(n)]


def same_as_single(got, single, precision, B):
    """Automatic tuning: the batched and single launches share the tile code."""
    return np.array_equal(got, single)


def make_blocks(shapes, seed):
    import torch
    g = torch.Generator(device="cuda").manual_seed(seed)
    rng = np.random.default_rng(seed)
    blocks = []
    for (A, F, T, B) in shapes:
        v = torch.randn((A, F, T, 2, 2), device="cuda", generator=g) * 16.0
        x = torch.clamp(torch.round(v), -127, 127).to(torch.int8)
        del v
        tau = rng.uniform(-50e-9, 50e-9, size=(B, A))
        w = O.phasor_ata(tau, np.ones((A, F, 1, 2)), 1.0e9, 0.5e6, 0.0, 0).astype(np.complex64)
        y = torch.empty((B, F, T, 2), dtype=torch.complex64, device="cuda")
        blocks.append((x, torch.from_numpy(w).cuda(), y))
    return blocks


@pytest.mark.parametrize("precision", ["f32", "f16x2"])
def test_batched_matches_single_launches_and_oracle(precision):
    from blade_amd import ops
    shapes = block_shapes(12, seed=3, Fs=(64, 128, 256), Ts=(1024, 2048, 4096))
    blocks = make_blocks(shapes, seed=4)
    bb = ops.BatchedBeamformer(blocks, precision)
    bb()
    for k, (x, w, y) in enumerate(blocks):
        single = ops.beamformer(x, w, False, False, precision)
        got = host(y)
        assert same_as_single(got, host(single), precision, shapes[k][3]), f"block {k} {shapes[k]}"
        fs = [0, x.shape[1] - 1]
        ref = O.beamformer_ata(host(x[:, fs]), host(w[:, :, fs]))
        assert O.norm_rel_err(got[:, fs], ref) < TOL, f"block {k} {shapes[k]}"


@pytest.mark.parametrize("precision", ["f32", "f16x2"])
def test_batched_full_configuration_sampled(precision):
    # The seeded 32-block workload at full size (up to F=1024, T=8192, B=64);
    # two channels of each block against the oracle.
    from blade_amd import ops
    shapes = block_shapes(32, seed=2026)
    blocks = make_blocks(shapes, seed=5)
    bb = ops.BatchedBeamformer(blocks, precision)
    bb()
    for k, (x, w, y) in enumerate(blocks):
        fs = [k % x.shape[1], x.shape[1] - 1]
        ref = O.beamformer_ata(host(x[:, fs]), host(w[:, :, fs]))
        assert O.norm_rel_err(host(y[:, fs]), ref) < TOL, f"block {k} {shapes[k]}"


def test_batched_tunings_agree():
    from blade_amd import ops
    shapes = block_shapes(6, seed=7, Fs=(64, 128), Ts=(1024, 2048))
    blocks = make_blocks(shapes, seed=8)
    bb = ops.BatchedBeamformer(blocks, "f32")
    bb()
    base = [host(y).copy() for (_, _, y) in blocks]
    for tuning in ({"chunks_per_item": 1, "workgroups_per_cu": 2},
                   {"chunks_per_item": 8, "workgroups_per_cu": 4},
                   {"beam_tiles": 2, "chunks_per_item": 2}):
        bb.set_tuning(tuning)
        bb()
        for k, (_, _, y) in enumerate(blocks):
            assert O.norm_rel_err(host(y), base[k]) < 1e-6, (tuning, k)


def test_batched_torch_graph_replays():
    import torch
    from blade_amd import ops
    shapes = block_shapes(5, seed=9, Fs=(64, 128), Ts=(1024, 2048))
    blocks = make_blocks(shapes, seed=10)
    bb = ops.BatchedBeamformer(blocks, "f16x2")
    bb()
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        bb()
    gen = torch.Generator(device="cuda").manual_seed(11)
    for rep in range(4):
        for (x, _, _) in blocks:        # new voltages, eager, between replays
            x.copy_(torch.randint(-127, 128, x.shape, dtype=torch.int8, device="cuda", generator=gen))
        graph.replay()
        torch.cuda.synchronize()
        for k, (x, w, y) in enumerate(blocks):
            assert np.array_equal(host(y), host(ops.beamformer(x, w, False, False, "f16x2"))), (rep, k)


def test_runner_pipeline_graph_capture():
    import blade_amd as bl
    from blade_amd import ops
    import torch

    shapes = block_shapes(4, seed=12, Fs=(64, 128), Ts=(1024, 2048))

    @bl.runner
    class BatchedPipeline:
        def __init__(self):
            self.input.bufs = [bl.array_tensor((A, F, T, 2), dtype=bl.ci8) for (A, F, T, B) in shapes]
            self.input.phasors = [bl.phasor_tensor((B, A, F, 1, 2), dtype=bl.cf32) for (A, F, T, B) in shapes]
            self.module.bf = bl.module(bl.batchbeamformer, {"enable_split_fp16": False},
                                       (self.input.bufs, self.input.phasors), it=bl.ci8, ot=bl.cf32)
            self.runner.enable_graph_capture(True)

        def transfer_in(self, bufs, phasors):
            for d, s in zip(self.input.bufs, bufs):
                self.copy(d, s)
            for d, s in zip(self.input.phasors, phasors):
                self.copy(d, s)

        def transfer_out(self, outs):
            for d, s in zip(outs, self.module.bf.get_output()):
                self.copy(d, s)

    p = BatchedPipeline()
    rng = np.random.default_rng(13)
    for rep in range(3):
        hin, hph, hout, refs = [], [], [], []
        for (A, F, T, B) in shapes:
            xi = bl.array_tensor((A, F, T, 2), dtype=bl.ci8, device=bl.cpu)
            xi.as_numpy()[...] = O.ci8_gaussian((A, F, T, 2), rng=rng)
            wi = bl.phasor_tensor((B, A, F, 1, 2), dtype=bl.cf32, device=bl.cpu)
            tau = rng.uniform(-50e-9, 50e-9, size=(B, A))
            wi.as_numpy()[...] = O.phasor_ata(tau, np.ones((A, F, 1, 2)), 1e9, 0.5e6, 0.0, 0)
            yo = bl.array_tensor((B, F, T, 2), dtype=bl.cf32, device=bl.cpu)
            hin.append(xi), hph.append(wi), hout.append(yo)
            refs.append(host(ops.beamformer(torch.from_numpy(xi.as_numpy().copy()).cuda(),
                                            torch.from_numpy(wi.as_numpy().copy()).cuda(), False, False, "f32")))
        assert p(hin, hph, hout)
        for k in range(len(shapes)):
            assert same_as_single(hout[k].as_numpy(), refs[k], "f32", shapes[k][3]), (rep, k)


def test_autotuner_picks_a_candidate(tmp_path, monkeypatch):
    from blade_amd import autotune, ops
    monkeypatch.setenv("BLADE_AMD_CACHE", str(tmp_path))
    shapes = block_shapes(6, seed=14, Fs=(64, 128), Ts=(1024, 2048))
    blocks = make_blocks(shapes, seed=15)
    tuning, ms = autotune.tune_batched(blocks, "f32", iters=2, warmup=1)
    assert tuning in list(autotune.candidates())
    assert ms > 0
    assert autotune.cached_tuning(blocks, "f32") == tuning
    bb = ops.BatchedBeamformer(blocks, "f32", tuning)
    bb()
    x, w, y = blocks[0]
    fs = [0]
    ref = O.beamformer_ata(host(x[:, fs]), host(w[:, :, fs]))
    assert O.norm_rel_err(host(y[:, fs]), ref) < TOL
