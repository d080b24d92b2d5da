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


# Host-only sanitizer builds of sanitize/host_concurrency.cc with ROCm's
# clang++ as a plain x86-64 C++ compiler (no HIP language mode, no device
# code; g++ 11 lacks the _Float16 that the tensor types use), the HIP header
# only for declarations.
CLANGXX = "/opt/rocm/llvm/bin/clang++"
SANITIZERS = {"tsan": ["-fsanitize=thread"], "asan": ["-fsanitize=address,undefined", "-fno-sanitize-recover=all"]}


def build_sanitizers(verbose=False):
    out = HERE / "_build"
    out.mkdir(exist_ok=True)
    src = HERE / "sanitize" / "host_concurrency.cc"
    for name, flags in SANITIZERS.items():
        exe = out / f"host_concurrency_{name}"
        cmd = [CLANGXX, "-std=c++20", "-O1", "-g", "-fno-omit-frame-pointer"] + flags + [
            "-D__HIP_PLATFORM_AMD__", f"-I{ROOT / 'include'}", f"-I{ROOT / 'blade_amd' / 'host'}",
            "-I/opt/rocm/include", str(src), "-o", str(exe), "-pthread",
            "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"]
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"building {exe.name} failed:\n{r.stdout}")
        if verbose:
            print(f"built {exe}")


if __name__ == "__main__":
    build(verbose=True)
    build_sanitizers(verbose=True)
    sys.exit(0)
