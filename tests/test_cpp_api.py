"""The typed C++ Flow API (include/graphflow.hpp) over the C ABI: host-only checks on the CPU, the
particle feedback network on the GPU against the oracle's scalar model."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "_build", "flow_example")


def example():
    if not os.path.exists(BIN):
        subprocess.run(["sh", os.path.join(HERE, "cpp", "build.sh")], check=True)
    return BIN


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
