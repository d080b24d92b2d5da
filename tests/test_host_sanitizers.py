"""Host code under sanitizers, and tensor ownership semantics, on the CPU.

* ``tests/cpp/_build/vector_semantics --capture``: the reference's Vector
  cases (tests/memory/vector/reference.cc:15-208) with the allocation
  profiler capturing, so nothing is allocated and no GPU is needed.
* ``tests/cpp/_build/host_concurrency_{tsan,asan}``: the tensor reference
  count under copies from 8 threads, the GUPPI reader's thread pool and
  positioned reads split over threads, and its header parser on malformed
  input, built with ThreadSanitizer and with AddressSanitizer + UBSan
  (host code only).
"""

import subprocess
from pathlib import Path

import pytest

BUILD = Path(__file__).resolve().parent / "cpp" / "_build"


def run(exe, *args):
    path = BUILD / exe
    assert path.exists(), f"tests/cpp/_build/{exe} is missing: run the build step"
    return subprocess.run([str(path), *map(str, args)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                          text=True, timeout=300)


def test_vector_semantics_without_allocation():
    r = run("vector_semantics", "--capture")
    assert r.returncode == 0, r.stdout
    assert "vector semantics: 0 failures" in r.stdout
    for case in ("empty creation", "zombie creation", "copy constructor", "move constructor",
                 "copy assignment", "move assignment", "initializer struct member"):
        assert case in r.stdout


@pytest.mark.parametrize("flavour,report", [("tsan", "WARNING: ThreadSanitizer"),
                                            ("asan", "ERROR: AddressSanitizer")])
def test_host_concurrency_under_sanitizer(tmp_path, flavour, report):
    r = run(f"host_concurrency_{flavour}", tmp_path)
    assert report not in r.stdout, r.stdout
    assert "runtime error" not in r.stdout, r.stdout       # UBSan
    assert r.returncode == 0, r.stdout
    assert "host concurrency: 0 failures" in r.stdout
