"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU).

* The oracle (oracle/bfvm_oracle.cpp) driven through every stage of its C ABI
  by tests/sanitize/oracle_driver.cpp: 1-3-D meshes, faces and interior
  Dirichlet voxels, a bulk, shared and inactive cells, internalized tracking,
  releases, recomputes, the FP32 solver and an out-of-grid cell.
* The boundary library (physicore_amd/host: mesh, agent storage,
  microenvironment, solver registry, VTK serializer, hip_solver) with the
  ported reference tests that run without a GPU (tests/cpp/test_hip_solver.cpp
  --host-only).

Both are built into a temporary directory with -fsanitize=address,undefined
and -fno-sanitize-recover (any report fails the run), leak detection on.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CXX = os.environ.get("CXX", "g++")
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer", "-g", "-O1"]
ENV = {**os.environ, "ASAN_OPTIONS": "detect_leaks=1:abort_on_error=0", "UBSAN_OPTIONS": "print_stacktrace=1",
       "OMP_NUM_THREADS": "2"}


def _build(cmd):
    if shutil.which(CXX) is None:
        pytest.skip(f"{CXX} not found")
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-3000:]


def _run(exe, *args):
    r = subprocess.run([exe, *args], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=600,
                       env=ENV)
    print(r.stdout[-3000:])
    assert "ERROR: AddressSanitizer" not in r.stdout and "runtime error:" not in r.stdout, r.stdout[-3000:]
    assert "ERROR: LeakSanitizer" not in r.stdout, r.stdout[-3000:]
    assert r.returncode == 0, r.stdout[-3000:]
    return r.stdout


def _instrumented(exe):
    r = subprocess.run(["nm", exe], stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True)
    return "__asan_init" in r.stdout or "__asan_report_load" in r.stdout


def test_oracle_under_sanitizers(tmp_path):
    exe = str(tmp_path / "oracle_driver")
    _build([CXX, "-std=c++20", "-fopenmp", "-ffp-contract=off", *SAN, "-I", os.path.join(ROOT, "oracle"),
            os.path.join(ROOT, "tests", "sanitize", "oracle_driver.cpp"), os.path.join(ROOT, "oracle", "bfvm_oracle.cpp"),
            "-o", exe])
    assert _instrumented(exe)
    out = _run(exe)
    assert "oracle driver: 0 failed" in out


def test_host_library_under_sanitizers(tmp_path):
    from physicore_amd import build as b

    lib = os.path.join(b.LIB_DIR, "libbfvm_hip.so")
    if not os.path.exists(lib):
        pytest.fail(f"{lib} is not built; run `python -m physicore_amd.build`")
    exe = str(tmp_path / "test_hip_solver_san")
    srcs = [os.path.join(b.HOST, f) for f in b.HOST_SRCS] + [os.path.join(ROOT, "tests", "cpp", "test_hip_solver.cpp")]
    inc = [x for d in b.HOST_INC for x in ("-I", d)]
    _build([CXX, "-std=c++20", "-pthread", *SAN, *inc, *srcs, "-L", b.LIB_DIR, "-lbfvm_hip",
            f"-Wl,-rpath,{b.LIB_DIR}", "-o", exe])
    assert _instrumented(exe)
    out = _run(exe, "--host-only")
    assert " 0 failed" in out
