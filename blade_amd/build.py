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
        deps = [src] + list(CSRC.glob("*.hh")) + list(INCLUDE.glob("*.h"))
        if all(obj.stat().st_mtime >= d.stat().st_mtime for d in deps):
            return obj
    cmd = [HIPCC] + HIP_FLAGS + ["-c", str(src), "-o", str(obj)]
    if src.suffix == ".cc":
        cmd = [HIPCC, "-O3", "-std=c++17", "-fPIC", f"-I{INCLUDE}", f"-I{CSRC}",
               "-c", str(src), "-o", str(obj)]
    _run(cmd)
    return obj


def build(incremental: bool = False, jobs: int | None = None, verbose: bool = True) -> Path:
    if not shutil.which(HIPCC) and not Path(HIPCC).exists():
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    OBJ.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cc"))
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, incremental), sources))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", str(tmp)]
         + [str(o) for o in objs] + ["-Wl,--no-undefined", "-lamdhip64"])
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(sources)} sources")
    return LIB


if __name__ == "__main__":
    build(incremental="--incremental" in sys.argv)
