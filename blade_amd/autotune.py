"""Autotuner for the batched (persistent) beamformer launch.

For a list of block shapes it times each candidate ``blade_batch_tuning``
on the device with HIP events (after a warm-up), keeps the fastest, and
caches it in a JSON file keyed by the precision and the multiset of block
shapes, so later runs of the same workload start from the tuned shape.

Candidates span the knobs of the launch:

* ``beam_tiles``: beams per wavefront (0 = 8 for <= 8 beams, else 16;
  1 = always 8, as (beam, re/im) row pairs; 2 = always 16, where FP32 uses
  the three-product form);
* ``chunks_per_item``: 256-sample time chunks per work item (load balance
  against per-item phasor staging);
* ``workgroups_per_cu``: workgroups per compute unit in the grid (2-4 keep
  every workgroup resident and walking items; 8-64 give short-lived
  workgroups the hardware refills as they finish, up to one per item).

The cache lives in ``$BLADE_AMD_CACHE`` (a directory) or
``~/.cache/blade_amd``; an unwritable location only disables the cache.
"""

from __future__ import annotations

import hashlib
import itertools
import json
import os
from pathlib import Path

import torch

from . import ops

__all__ = ["candidates", "signature", "tune_batched", "cached_tuning"]


def candidates():
    for bt, cpi, wpc in itertools.product((0, 1, 2), (0, 1, 2, 4, 8), (2, 3, 4, 8, 16, 64)):
        yield {"time_samples_per_lane": 4, "beam_tiles": bt, "chunks_per_item": cpi,
               "workgroups_per_cu": wpc}


def signature(shapes, precision):
    key = json.dumps({"precision": precision, "shapes": sorted(list(map(list, shapes)))})
    return hashlib.sha1(key.encode()).hexdigest()


def _cache_path():
    d = os.environ.get("BLADE_AMD_CACHE") or os.path.join(Path.home(), ".cache", "blade_amd")
    return Path(d) / "autotune.json"


def _load_cache():
    try:
        with open(_cache_path()) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def _store_cache(cache):
    p = _cache_path()
    try:
        p.parent.mkdir(parents=True, exist_ok=True)
        tmp = p.with_suffix(".tmp")
        with open(tmp, "w") as fh:
            json.dump(cache, fh, indent=1)
        os.replace(tmp, p)
    except OSError:
        pass


def _shapes(blocks):
    out = []
    for (x, w, _) in blocks:
        A, F, T = x.shape[0], x.shape[1], x.shape[2]
        out.append((int(A), int(F), int(T), int(w.shape[0])))
    return out


def cached_tuning(blocks, precision="f32"):
    """The cached best tuning for this workload, or None."""
    entry = _load_cache().get(signature(_shapes(blocks), precision))
    return entry["tuning"] if entry else None


def _time(bb, iters, stream):
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(iters):
        bb(stream)
    e1.record(stream)
    e1.synchronize()
    return e0.elapsed_time(e1) / iters


def tune_batched(blocks, precision="f32", iters=5, warmup=2, use_cache=True, verbose=False):
    """Best ``blade_batch_tuning`` (as a dict) for these blocks, and its time in ms.

    ``blocks``: list of ``(x, phasors, out)`` device tensors as for
    :class:`blade_amd.ops.BatchedBeamformer` (their outputs are overwritten).
    """
    sig = signature(_shapes(blocks), precision)
    cache = _load_cache() if use_cache else {}
    if use_cache and sig in cache:
        return cache[sig]["tuning"], cache[sig]["ms"]
    stream = torch.cuda.current_stream()
    bb = ops.BatchedBeamformer(blocks, precision)
    # warm the clocks with the automatic shape
    for _ in range(warmup):
        _time(bb, iters, stream)
    best, best_ms = None, float("inf")
    for cand in candidates():
        bb.set_tuning(cand)
        bb(stream)
        ms = min(_time(bb, iters, stream) for _ in range(2))
        if verbose:
            print(f"  {cand}: {ms:.3f} ms", flush=True)
        if ms < best_ms:
            best, best_ms = dict(cand), ms
    if use_cache:
        cache = _load_cache()
        cache[sig] = {"tuning": best, "ms": best_ms, "shapes": _shapes(blocks), "precision": precision}
        _store_cache(cache)
    return best, best_ms
