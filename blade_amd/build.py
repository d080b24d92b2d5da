"""Build recipe for the gfx950 library ``blade_amd/libblade_amd.so``.

Every ``csrc/*.hip`` file is compiled by hipcc for gfx950 only and linked,
together with ``csrc/*.cc``, into one shared library exporting the C-ABI of
``include/blade_amd.h``.  ``python -m blade_amd.build [--incremental]``.
"""

from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"
OBJ = PKG / "_build"
LIB = PKG / "libblade_amd.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# Packed-FP32 instructions are disabled for every HIP translation unit.
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wno-unused-result",
    f"-I{INCLUDE}", f"-I{CSRC}",
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops' is not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    return out


def _compile(src: Path, incremental: bool) -> Path:
    obj = OBJ / (src.name + ".o")
    if incremental and obj.exists():
        deps = [src] + list(CSRC.glob("*.hh")) + list((PKG / "host").glob("*.hh")) + list(INCLUDE.rglob("*.h*"))
        if all(obj.stat().st_mtime >= d.stat().st_mtime for d in deps):
            return obj
    cmd = [HIPCC] + HIP_FLAGS + ["-c", str(src), "-o", str(obj)]
    if src.suffix == ".cc":
        # Host-only C++ (runtime, Module/Pipeline/Runner): no device code.
        cmd = [HIPCC, "-O2", "-std=c++20", "-fPIC", "-fvisibility=default",
               "-Wno-unused-result", f"-I{INCLUDE}", f"-I{CSRC}",
               "-c", str(src), "-o", str(obj)]
    _run(cmd)
    return obj


def _compile_bindings(incremental: bool) -> Path:
    import sysconfig

    import pybind11
    src = PKG / "host" / "bindings.cc"
    obj = OBJ / "bindings.cc.o"
    if incremental and obj.exists():
        deps = [src] + list(INCLUDE.rglob("*.h*"))
        if all(obj.stat().st_mtime >= d.stat().st_mtime for d in deps):
            return obj
    _run([HIPCC, "-O1", "-std=c++20", "-fPIC", "-fvisibility=hidden", "-Wno-unused-result",
          f"-I{INCLUDE}", f"-I{pybind11.get_include()}", f"-I{sysconfig.get_paths()['include']}",
          "-c", str(src), "-o", str(obj)])
    return obj


def build(incremental: bool = False, jobs: int | None = None, verbose: bool = True) -> Path:
    import sysconfig
    if not shutil.which(HIPCC) and not Path(HIPCC).exists():
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    OBJ.mkdir(exist_ok=True)
    sources = (sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cc"))
               + sorted(s for s in (PKG / "host").glob("*.cc") if s.name != "bindings.cc"))
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        fut_b = ex.submit(_compile_bindings, incremental)   # the slowest unit starts first
        objs = list(ex.map(lambda s: _compile(s, incremental), sources))
        bind_obj = fut_b.result()
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", str(tmp)]
         + [str(o) for o in objs] + ["-Wl,--no-undefined", "-lamdhip64", "-lhiprtc",
                                      "-lrocprofiler-sdk-roctx"])
    os.replace(tmp, LIB)
    # pybind11 extension (the Python mirror of the C++ API) over the library.
    ext = PKG / ("_blade_ext" + sysconfig.get_config_var("EXT_SUFFIX"))
    tmp = ext.with_suffix(".so.tmp")
    _run([HIPCC, "-shared", "-fPIC", str(bind_obj), "-o", str(tmp), f"-L{PKG}", "-lblade_amd",
          "-Wl,-rpath,$ORIGIN", "-lamdhip64"])
    os.replace(tmp, ext)
    if verbose:
        print(f"built {LIB} from {len(sources)} sources and {ext.name}")
    return LIB


if __name__ == "__main__":
    build(incremental="--incremental" in sys.argv)
