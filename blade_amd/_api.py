"""Python mirror of the reference package ``blade`` over the ``_blade_ext``
extension (imported by ``blade_amd/__init__.py`` once the extension loads)."""

from __future__ import annotations

import importlib
import inspect as _inspect
import sys as _sys
from types import SimpleNamespace

from . import _blade_ext as _ext

# Register every submodule under the reference's extension names
# (blade._beamformer_impl, blade._mem_impl, ...).
for _name in dir(_ext):
    if _name.startswith("_") and _name.endswith("_impl"):
        _sys.modules[f"{__package__}.{_name}"] = getattr(_ext, _name)

_const = _ext._const_impl
_hidden = _ext._hidden_impl
_mem = _ext._mem_impl

result = _const.result
pol = _const.pol
stream = _const.stream
xyz = _const.xyz
uvw = _const.uvw
lla = _const.lla
ra_dec = _const.ra_dec
set_device = _const.set_device


class _Constant:
    """Named constant (reference python/blade/_internal/__init__.py:56-66)."""

    def __init__(self, value):
        self._value = value

    def __repr__(self):
        return f"Constant(value={self._value})"

    @property
    def value(self):
        return self._value


_telescope_lst = ["ata", "vla", "meerkat", "generic"]
_device_lst = ["cpu", "cuda", "unified"]
_types_lst = ["f16", "f32", "f64", "i8", "i16", "i32", "i64", "u8", "u16", "u32", "u64",
              "cf16", "cf32", "cf64", "ci8", "ci16", "ci32", "ci64", "cu8", "cu16", "cu32", "cu64"]
_modules_lst = ["beamformer", "cast", "channelizer", "detector", "duplicate", "gather", "guppi_reader",
                "guppi_writer", "permutation",
                "phasor", "polarizer", "modeh", "modeb", "batchbeamformer"]

for _n in _telescope_lst + _device_lst + _types_lst + _modules_lst:
    globals()[_n] = _Constant(_n)


def _fetch(name="runner"):
    """The pipeline a module or tensor is being created in: the `runner`
    attribute of the calling @bl.runner instance (reference _internal:_Fetch)."""
    frame = _inspect.currentframe()
    for _ in range(3):
        if frame is None:
            break
        frame = frame.f_back
    if frame is not None and "self" in frame.f_locals:
        inst = frame.f_locals["self"]
        if hasattr(inst, name):
            return getattr(inst, name)
    if name == "runner":
        return _fetch("pipeline")
    return None


def _is_duet(obj):
    return type(obj).__name__.endswith("_duet")


def _sanitize_duet(obj, index=0):
    def item(x):
        return x[index] if _is_duet(x) else x
    if isinstance(obj, tuple):
        return tuple(item(x) for x in obj)
    if isinstance(obj, dict):
        return {item(k): item(v) for k, v in obj.items()}
    return item(obj)


# ----------------------------------------------------------------- tensors

def _create_array(kind, shape, dtype, device):
    pipeline = _fetch()
    kind = kind + "_duet" if pipeline else kind
    if isinstance(shape, int):
        shape = (shape,)
    unified = device is globals()["unified"] or getattr(device, "value", None) == "unified"
    dev = "cuda" if unified else device.value
    cls = getattr(getattr(getattr(_mem, dev), dtype.value), kind)
    return cls(tuple(shape), unified=unified)


def array_tensor(shape, dtype=None, device=None):
    """[A, F, T, P] tensor (reference python/blade/_memory/__init__.py:18-31)."""
    return _create_array("array_tensor", shape, dtype or globals()["f32"], device or globals()["cuda"])


def phasor_tensor(shape, dtype=None, device=None):
    """[B, A, F, T, P] tensor."""
    return _create_array("phasor_tensor", shape, dtype or globals()["f32"], device or globals()["cuda"])


def delay_tensor(shape, dtype=None, device=None):
    """[B, A] tensor."""
    return _create_array("delay_tensor", shape, dtype or globals()["f32"], device or globals()["cuda"])


def tensor(shape, dtype=None, device=None):
    """1-D tensor."""
    return _create_array("tensor", shape, dtype or globals()["f32"], device or globals()["cuda"])


def _resolve(t, slot):
    return t[slot] if _is_duet(t) else t


def _dims(shape):
    return tuple(shape[i] for i in range(len(shape)))


def _copy(dst, src, stream_handle=0, sync=True):
    if dst.size_bytes() != src.size_bytes():
        raise ValueError(f"Size mismatch between source ({src.size_bytes()} B) and destination "
                         f"({dst.size_bytes()} B).")
    if _dims(dst.shape) != _dims(src.shape):
        raise ValueError(f"Shape mismatch between source ({src.shape}) and destination ({dst.shape}).")
    return _ext._copy_impl.copy_bytes(dst.data_ptr(), src.data_ptr(), dst.size_bytes(),
                                      stream_handle, sync)


def copy(dst, src):
    """Synchronous copy between two tensors (reference python/blade/ext_copy.cc:16-18:
    destination first)."""
    return _copy(_resolve(dst, 0), _resolve(src, 0))


# ----------------------------------------------------------------- modules

def module(name, config, input, it=None, ot=None, telescope=None):
    """Instantiate a module (reference python/blade/_modules/__init__.py:17-160)."""
    it = it or globals()["cf32"]
    ot = ot or globals()["cf32"]
    telescope = telescope or globals()["generic"]
    for arg, what in ((name, "name"), (it, "input data type"), (ot, "output data type"),
                      (telescope, "telescope")):
        if not isinstance(arg, _Constant):
            raise ValueError(f"Module {what} has to be a constant type.")
    ext_name = name.value.split("_")[0]
    try:
        caller = importlib.import_module(f"{__package__}._{ext_name}_impl")
    except ModuleNotFoundError:
        raise ModuleNotFoundError(f"Can't find specified module extension ({ext_name}).")
    if any(s.startswith("tel_") for s in dir(caller)):
        tels = [s[4:] for s in dir(caller) if s.startswith("tel_")]
        if telescope.value in tels:
            caller = getattr(caller, f"tel_{telescope.value}")
        elif telescope.value != "generic":
            raise AttributeError(f"The module '{name.value}' only supports these telescopes: {', '.join(tels)}.")
    elif telescope.value != "generic":
        raise AttributeError(f"The module '{name.value}' only supports generic telescopes (telescope=bl.generic).")
    parts = name.value.split("_")
    taint = parts[1] if len(parts) > 1 else None
    if any(s.startswith("taint_") for s in dir(caller)):
        taints = [s[6:] for s in dir(caller) if s.startswith("taint_")]
        if taint not in taints:
            raise AttributeError(f"The module '{name.value}' only supports these taints: {', '.join(taints)}.")
        caller = getattr(caller, f"taint_{taint}")
    elif taint:
        raise AttributeError(f"The module '{name.value}' doesn't support taints.")
    if any(s.startswith("in_") for s in dir(caller)):
        ins = [s[3:] for s in dir(caller) if s.startswith("in_")]
        if it.value not in ins:
            raise AttributeError(f"The module '{name.value}' only supports these input types: {', '.join(ins)}.")
        caller = getattr(caller, f"in_{it.value}")
    if any(s.startswith("out_") for s in dir(caller)):
        outs = [s[4:] for s in dir(caller) if s.startswith("out_")]
        if ot.value not in outs:
            raise AttributeError(f"The module '{name.value}' with input '{it.value}' only supports these "
                                 f"output types: {', '.join(outs)}.")
        caller = getattr(caller, f"out_{ot.value}")
    if "mod" not in dir(caller):
        raise RuntimeError(f"The module '{name.value}' doesn't have a valid module handle.")
    mod_cls = caller.mod
    cfg_cls = caller.config
    in_cls = caller.input
    if isinstance(config, tuple):
        cfg = cfg_cls(*config)
    elif isinstance(config, dict):
        cfg = cfg_cls(**config)
    elif isinstance(config, int):
        cfg = cfg_cls(config)
    else:
        raise ValueError("Config should be a Tuple, Dict, or Int.")
    sanitized = _sanitize_duet(input)
    if isinstance(input, tuple):
        inp = in_cls(*sanitized)
    elif isinstance(input, dict):
        inp = in_cls(**sanitized)
    else:
        inp = in_cls(sanitized)
    pipeline = _fetch()
    if pipeline:
        if pipeline.commited():
            raise RuntimeError("Can't connect new module after Pipeline is commited.")
        inst = mod_cls(cfg, inp, pipeline.stream())
        if isinstance(inst, _hidden.bundle):
            for m in inst.modules():
                pipeline.add_module(m)
            # the bundle object owns its modules' inputs
            if not hasattr(pipeline, "_keep_bundles"):
                pipeline._keep_bundles = []
            pipeline._keep_bundles.append(inst)
        else:
            pipeline.add_module(inst)
    else:
        inst = mod_cls(cfg, inp, stream())
    return inst


# ------------------------------------------------------------ runner/pipeline

def pipeline():
    return _ext._pipeline_impl.pipeline()


def runner(cls):
    """Decorator turning a class into a synchronous pipeline
    (reference python/blade/_runner/__init__.py:6-86)."""

    class BasePipeline(cls):
        def __init__(self, *args, **kwargs):
            self.runner = _ext._runner_impl.runner()
            self.input = SimpleNamespace()
            self.output = SimpleNamespace()
            self.module = SimpleNamespace()
            super().__init__(*args, **kwargs)

        def copy(self, dst, src):
            """Runner.copy: duets resolve to the head slot, on the head stream."""
            head = self.runner.head()
            res = _copy(_resolve(dst, head), _resolve(src, head),
                        self.runner.stream_handle(head), sync=False)
            assert res == result.success

        def enqueue(self, input_callback, output_callback, input_id, output_id):
            res = self.runner.enqueue(input_callback, output_callback, input_id, output_id)
            if res in (result.success, result.runner_queue_full, result.runner_queue_empty,
                       result.runner_queue_none_available):
                return res
            raise RuntimeError(f"enqueue failed: {res}")

        def dequeue(self, callback):
            res = self.runner.dequeue(callback)
            if res in (result.success, result.runner_queue_full, result.runner_queue_empty,
                       result.runner_queue_none_available):
                return res
            raise RuntimeError(f"dequeue failed: {res}")

        def __call__(self, *args):
            if not hasattr(self, "transfer_in"):
                raise AttributeError("The 'transfer_in' function is missing in the pipeline.")
            if not hasattr(self, "transfer_out"):
                raise AttributeError("The 'transfer_out' function is missing in the pipeline.")
            for ns in (self.input, self.output):
                for value in vars(ns).values():
                    if _is_duet(value):
                        value.set(0)
            n_in = self.transfer_in.__code__.co_argcount - 1
            n_out = self.transfer_out.__code__.co_argcount - 1
            if len(args) != n_in + n_out:
                raise ValueError(f"Expected {n_in + n_out} arguments, but got {len(args)}.")
            will_output = self.runner.will_output()
            self.transfer_in(*args[:n_in])
            res = self.runner.compute(0)
            assert res == result.success, res
            if will_output:
                self.transfer_out(*args[n_in:])
            assert self.runner.synchronize(0) == result.success
            return will_output

        def synchronize(self):
            assert self.runner.synchronize(0) == result.success

    BasePipeline.__name__ = cls.__name__
    BasePipeline.__qualname__ = cls.__qualname__
    return BasePipeline


__all__ = _telescope_lst + _device_lst + _types_lst + _modules_lst + ["array_tensor", "phasor_tensor", "delay_tensor", "tensor", "copy", "module", "runner",
           "pipeline", "result", "pol", "stream", "xyz", "uvw", "lla", "ra_dec", "set_device"]
