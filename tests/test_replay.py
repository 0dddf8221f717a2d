"""The library's host code and its kernels replayed on the CPU (tests/replay:
device memory as host memory, one fiber per GPU thread, wavefront exchanges
in lockstep) under AddressSanitizer + UndefinedBehaviorSanitizer and
ThreadSanitizer.

* teardown: parents closed before their children (context before geometry,
  tree and vortex-particle case; geometry before its case).  Before the
  context and geometry held references, this scenario was a heap
  use-after-free: the cause of the glibc heap abort at interpreter exit seen
  after the vortex-particle case's GPU run.
* vp: the vortex-particle driver on the reference's steadyTest case against
  its stored steps, handles left to the interpreter (finalized together, in
  no fixed order) - the teardown that aborted.
* vpcfg: the configured form of the case (algebraic fixed-core kernel,
  particles split along each shedding panel, FP32 sums) on a small swept,
  tapered wing, and theta = 0 trees of that kernel against the oracle.
* panels: the panel -> point velocity and stretching sums (far pass, queued
  near pass) against the oracle.
* assembly: the far and near passes (tile culling, lane-permute compaction)
  against the FP64 oracle, 1e-12 of max |A| and max |B|.

Each scenario runs in a child process with the sanitizer runtime preloaded;
a sanitizer report aborts it.
"""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "replay"))

import build as replay_build  # noqa: E402

pytestmark = pytest.mark.skipif(not replay_build.available(), reason="g++ with ASan/TSan runtimes not available")


def run(mode, *args, timeout=900):
    lib = replay_build.build(mode)
    r = subprocess.run([sys.executable, os.path.join(HERE, "replay", "run_case.py"), lib] + list(args),
                       env=replay_build.preload_env(mode), stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=timeout)
    assert r.returncode == 0, r.stdout[-6000:]
    for bad in ("AddressSanitizer", "runtime error:", "ThreadSanitizer"):
        assert bad not in r.stdout, r.stdout[-6000:]
    return r.stdout


def test_teardown_in_any_order_asan():
    assert "ok teardown" in run("asan", "teardown")


def test_assembly_passes_match_oracle_asan():
    assert "ok assembly" in run("asan", "assembly")


def test_vp_steps_and_interpreter_teardown_asan():
    out = run("asan", "vp", "6")
    assert "ok vp step 6" in out and "ok vp finish" in out


def test_panel_sums_match_oracle_asan():
    """Panel -> point velocity and stretching: far pass and queued near pass
    against the oracle, FP64 and with the far field in FP32."""
    out = run("asan", "panels")
    assert out.count("ok panels") == 2


def test_vp_config_asan():
    """The configured form (algebraic core, two particles per panel, FP64 and
    FP32 sums) for 10 steps, then theta = 0 trees against the oracle."""
    assert "ok vp config" in run("asan", "vpcfg", "10")


def test_vp_steps_tsan():
    """The post-table build runs on a host thread beside the steps."""
    out = run("tsan", "vp", "5")
    assert "ok vp finish" in out
