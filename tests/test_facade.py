"""The C++ drop-in facade (atlas-engine_amd/facade/atlas_rt_volume.h) keeps
the engine's Volume::BVH call sites source-compatible; the compiled check
program (built by the package Makefile) runs host-only on the CPU and
traces through the C ABI on the GPU."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "atlas-engine_amd", "_build", "facade_test")


def _run(env_extra):
    assert os.path.exists(BIN), "facade_test not built (run __graft_entry__.build())"
    env = dict(os.environ, **env_extra)
    p = subprocess.run([BIN], capture_output=True, text=True, env=env, timeout=120)
    assert p.returncode == 0, f"facade_test exit {p.returncode}: {p.stdout} {p.stderr}"
    assert "facade ok" in p.stdout
    return p.stdout


def test_facade_host_only():
    out = _run({"HIP_VISIBLE_DEVICES": "-1", "ROCR_VISIBLE_DEVICES": "-1"})
    assert "host only" in out


@pytest.mark.gpu
def test_facade_traces_on_gpu(torch_cuda):
    out = _run({})
    assert "hits" in out and "host only" not in out
