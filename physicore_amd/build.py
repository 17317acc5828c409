"""Build recipe of the native parts, run in-tree (no JIT cache):

  physicore_amd/lib/libbfvm_hip.so      HIP kernels + C ABI (hipcc, gfx950)
  physicore_amd/lib/libbfvm_host.so     C++20 hip_solver + biofvm boundary (g++)
  oracle/liboracle.so                   CPU oracle (test infrastructure, make)

`python -m physicore_amd.build` or physicore_amd.build.build_all().
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_DIR = os.path.join(HERE, "lib")
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PHYSICORE_AMD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++20", "-fPIC", "-shared",
    # no packed FP32 instructions in any kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wall",
]


def run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(line for line in r.stdout.splitlines() if "packed-fp32-ops" not in line)
    if out.strip():
        print(out, flush=True)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}")


HIP_SOURCES = ([f"bfvm_twist_{a}{h}.hip" for a in "yz" for h in "abcde"] +
               [f"bfvm_sweeps_xt{s}{h}.hip" for s in (2, 4, 8) for h in "ab"] +
               [f"bfvm_sweeps_xg{s}{h}.hip" for s in range(1, 9) for h in "ab"] +
               [f"bfvm_sweeps_xgt{s}.hip" for s in range(1, 9)] +
               [f"bfvm_pipe_x{s}.hip" for s in (1, 2, 4, 8)] + [f"bfvm_pipe_{a}{g}.hip" for a in "yz" for g in range(6)] +
               [f"bfvm_sweeps_{a}64{h}.hip" for a in "xyz" for h in "ab"] +
               ["bfvm_sweeps_x32.hip", "bfvm_sweeps_y32.hip", "bfvm_sweeps_z32.hip",
                "bfvm_sweeps_rows.hip", "bfvm_sweeps_xscan.hip", "bfvm_sweeps.hip", "bfvm_kernels.hip",
                "bfvm_context.hip"])


def build_hip(jobs=None):
    """Compiles the HIP translation units in parallel (the unrolled sweep
    instantiations dominate the build time) and links libbfvm_hip.so."""
    from concurrent.futures import ThreadPoolExecutor

    os.makedirs(LIB_DIR, exist_ok=True)
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    compile_flags = [f for f in HIP_FLAGS if f != "-shared"]
    inc = ["-I", os.path.join(ROOT, "include"), "-I", CSRC]

    def compile_one(src):
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        run([HIPCC, *compile_flags, *inc, "-c", os.path.join(CSRC, src), "-o", obj])
        return obj

    jobs = jobs or min(len(HIP_SOURCES), os.cpu_count() or 4, 16)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, HIP_SOURCES))
    out = os.path.join(LIB_DIR, "libbfvm_hip.so")
    run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


# the reference interface the backend plugs into (host/include, host/src) and
# the backend itself (host/hip_solver: the sources that drop into the
# reference tree as reactions-diffusion/biofvm/kernels/hip_solver)
HOST_SRCS = ["src/mesh.cpp", "src/agent_data.cpp", "src/microenvironment.cpp", "src/solver_registry.cpp",
             "src/vtk_serializer.cpp", "hip_solver/src/hip_solver.cpp"]
HOST_INC = [os.path.join(HOST, "include"), os.path.join(HOST, "hip_solver", "include"), os.path.join(ROOT, "include")]
CXX = os.environ.get("CXX", "g++")
HOST_FLAGS = ["-std=c++20", "-O2", "-fPIC", "-pthread", "-Wall", "-Wextra", "-Wno-unused-parameter"]


def build_host():
    """C++20 hip_solver + the biofvm boundary it plugs into; talks to the GPU only
    through libbfvm_hip.so's C ABI (no HIP headers)."""
    os.makedirs(LIB_DIR, exist_ok=True)
    out = os.path.join(LIB_DIR, "libbfvm_host.so")
    srcs = [os.path.join(HOST, f) for f in HOST_SRCS]
    inc = [x for d in HOST_INC for x in ("-I", d)]
    run([CXX, *HOST_FLAGS, "-shared", *inc, *srcs, "-L", LIB_DIR, "-lbfvm_hip", "-Wl,-rpath,$ORIGIN", "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_cpp_tests():
    """The reference's solver tests ported onto hip_solver (tests/cpp), linked to libbfvm_host.so."""
    src = os.path.join(ROOT, "tests", "cpp", "test_hip_solver.cpp")
    if not os.path.exists(src):
        return None
    bindir = os.path.join(ROOT, "tests", "cpp", "bin")
    os.makedirs(bindir, exist_ok=True)
    out = os.path.join(bindir, "test_hip_solver")
    inc = [x for d in HOST_INC for x in ("-I", d)]
    run([CXX, *HOST_FLAGS, *inc, src, "-L", LIB_DIR, "-lbfvm_host", "-lbfvm_hip", f"-Wl,-rpath,{LIB_DIR}",
         "-Wl,-rpath,$ORIGIN/../../../physicore_amd/lib", "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_oracle():
    run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])


def build_all():
    build_hip()
    build_host()
    build_cpp_tests()
    build_oracle()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
