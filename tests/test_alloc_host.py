"""Host test of the device-memory cache's bookkeeping (cosmicscanner_amd/csrc/cs_alloc.cpp).

The allocator is compiled with g++ against a host-only stand-in of the HIP calls it makes
(tests/cpp/hip_host_stub) and the address and undefined-behaviour sanitizers, then driven by
several threads through random allocations, growth and releases inside and outside the deferred
and idle release sections, with and without a capped runtime (the out-of-memory retry).  Any use
of a released block, double release or leak fails the run.  Before that, the reuse rule of a
deferred section is checked directly: a block released inside it is not handed out again until
the section has ended (its device work may still run), and is handed out after.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def stress_bin(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    out = str(tmp_path_factory.mktemp("alloc") / "alloc_host")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
           "-I" + os.path.join(ROOT, "tests/cpp/hip_host_stub"), "-I" + os.path.join(ROOT, "cosmicscanner_amd/csrc"),
           os.path.join(ROOT, "tests/cpp/test_alloc_host.cpp"), os.path.join(ROOT, "cosmicscanner_amd/csrc/cs_alloc.cpp"),
           "-o", out, "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return out


@pytest.mark.parametrize("threads,rounds,capacity", [(4, 150, 0), (6, 150, 600_000_000)])
def test_device_cache_bookkeeping(stress_bin, threads, rounds, capacity):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1", UBSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([stress_bin, str(threads), str(rounds), str(capacity)], capture_output=True, text=True,
                       env=env, timeout=300)
    assert r.returncode == 0, (r.stdout + r.stderr)[-4000:]
    assert "live bytes after release: 0" in r.stdout
    if capacity:
        assert "allocation failures: 0" not in r.stdout   # the capped run reached the retry path
