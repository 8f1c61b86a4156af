"""The typed C++ Flow API (include/graphflow.hpp) over the C ABI: host-only checks on the CPU, the
particle feedback network on the GPU against the oracle's scalar model."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "_build", "flow_example")


STAGE_BIN = os.path.join(HERE, "cpp", "_build", "stage_batches")


def example(binary=BIN):
    if not os.path.exists(binary):
        subprocess.run(["sh", os.path.join(HERE, "cpp", "build.sh")], check=True)
    return binary


def test_cpp_api_host_checks():
    out = subprocess.run([example(), "cpu"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "cpu checks ok" in out.stdout


@pytest.mark.gpu
def test_cpp_api_feedback_matches_oracle():
    import graphflow_amd as g
    import oracle as O

    N = 300_000
    out = subprocess.run([example(), "gpu", str(N)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    fields = dict(zip(out.stdout.split()[0::2], out.stdout.split()[1::2]))
    assert int(fields["sunk"]) == N
    w, _ = O.feedback_scalar(42, N, threads=16)
    assert int(fields["digest"]) == g.record_digest(w)
    assert "invalid_argument: yes" in out.stdout


@pytest.mark.gpu
def test_cpp_stage_batches_through_the_c_abi():
    """A graphFlow GraphRunner's HipStage binding, one partition batch at a time: launch / advance /
    finish and gf_transform_batch / advance / finish give three outputs of the whole chunk every
    time, the same outputs, and leave the pool whole (tests/cpp/stage_batches.cpp; it prints the
    time per batch)."""
    out = subprocess.run([example(STAGE_BIN), "500"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("ok batches 500"), out.stdout
    print(out.stdout.strip())
