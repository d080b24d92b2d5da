"""Builds the C++ API test programs of this folder against include/ and
blade_amd/libblade_amd.so into tests/cpp/_build/ (host code only)."""

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def build(verbose=False):
    out = HERE / "_build"
    out.mkdir(exist_ok=True)
    lib = ROOT / "blade_amd"
    for src in sorted(HERE.glob("*.cc")):
        exe = out / src.stem
        cmd = [HIPCC, "-O2", "-std=c++20", f"-I{ROOT / 'include'}", str(src), "-o", str(exe),
               f"-L{lib}", "-lblade_amd", f"-Wl,-rpath,{lib}", "-Wl,-rpath,$ORIGIN/../../../blade_amd"]
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"building {src.name} failed:\n{r.stdout}")
        if verbose:
            print(f"built {exe}")


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
