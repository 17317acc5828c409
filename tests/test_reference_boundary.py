"""The drop-in backend compiles against the reference's own headers.

hip_solver (physicore_amd/host/hip_solver) is meant to sit in the reference
tree as reactions-diffusion/biofvm/kernels/hip_solver and register through the
reference's solver registry (reactions-diffusion/biofvm/src/solver_registry.cpp:38-52).
This compiles its sources with the reference's include directories in place
of this repository's mirror of them, plus this library's C ABI header.  The
reference's CMake generates biofvm/biofvm_export.h; a one-line stand-in
defines its export macro here.  Skipped where the reference checkout is absent.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
REF_INC = [os.path.join(REF, "common", "include"), os.path.join(REF, "reactions-diffusion", "biofvm", "include")]
BACKEND = os.path.join(ROOT, "physicore_amd", "host", "hip_solver")

pytestmark = pytest.mark.skipif(not all(os.path.isdir(d) for d in REF_INC) or shutil.which("g++") is None,
                                reason="reference checkout or g++ absent")


def _compile(tmp_path, src, extra=()):
    exp = tmp_path / "biofvm"
    exp.mkdir(exist_ok=True)
    (exp / "biofvm_export.h").write_text("#pragma once\n#define BIOFVM_EXPORT\n")
    inc = [str(tmp_path), *REF_INC, os.path.join(ROOT, "include"), os.path.join(BACKEND, "include")]
    cmd = ["g++", "-std=c++20", "-fsyntax-only", "-Wall", *extra, *(x for d in inc for x in ("-I", d)), str(src)]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout


def test_backend_compiles_against_reference_headers(tmp_path):
    _compile(tmp_path, os.path.join(BACKEND, "src", "hip_solver.cpp"))


def test_registration_compiles_in_reference_tree(tmp_path):
    """What the reference's attachment point gains:
    attach_to_registry() called beside the OpenMP and Thrust backends, and the
    registry's factory handing out a hip_solver."""
    src = tmp_path / "attach.cpp"
    src.write_text(
        "#include <biofvm/solver_registry.h>\n"
        "#include <hip_solver/register_solver.h>\n"
        "#include <hip_solver/hip_solver.h>\n"
        "void attach() { physicore::biofvm::kernels::hip_solver::attach_to_registry(); }\n"
        "physicore::biofvm::solver_ptr make() {\n"
        "  return physicore::biofvm::solver_registry::instance().get(\"hip_solver\"); }\n")
    _compile(tmp_path, src)
