"""Runs the reference's solver tests ported onto the C++20 hip_solver
(tests/cpp/test_hip_solver.cpp): the host-side cases on the CPU, all cases on
the GPU."""
import os
import subprocess

import pytest

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "bin", "test_hip_solver")


def _run(*args, timeout=600):
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} is not built; run `python -m physicore_amd.build`")
    r = subprocess.run([BIN, *args], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=timeout)
    print(r.stdout[-4000:])
    return r


def test_host_cases():
    r = _run("--host-only")
    assert r.returncode == 0, r.stdout[-2000:]
    assert " 0 failed" in r.stdout


@pytest.mark.gpu
def test_all_cases_on_gpu():
    r = _run(timeout=900)
    assert r.returncode == 0, r.stdout[-3000:]
    assert " 0 failed" in r.stdout
