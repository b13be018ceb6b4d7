"""Builds the reference's `gw` executable with the hip backend wired in the way
the integration guide at the repository root (GUIDE below) tells a miniGW
maintainer to: its ```cpp file=...``` blocks are written into a scratch tree,
its ```diff``` blocks are applied to copies of the reference files they name,
and the result is compiled with the reference's own sources (from REF, where
they lie) and linked against minigw_amd/lib/libmgw_hip.so.  The scratch tree
lives in a temporary directory outside the repository; only the executable is
written (by default to oracle/_ref/gw_hip, git-ignored).

Test infrastructure: tests/test_integration.py links it on the CPU and
tests/test_gw_executable.py runs it on the GPU.  src/hdf5_input.cpp is
replaced by the guide's src/hdf5_input_hip.cpp, so no HDF5 headers or
libraries are needed; no MPI, BLAS or LAPACK either (the executable carries
the reference backend and the hip backend).

    python oracle/build_integration.py [--out PATH] [--ref DIR] [-j N]
"""
import argparse
import os
import re
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
GUIDE = os.path.join(REPO, "INTEGRATION.md")
LIB_DIR = os.path.join(REPO, "minigw_amd", "lib")
DEFAULT_OUT = os.path.join(HERE, "_ref", "gw_hip")

# Translation units of the reference's gw executable that build without
# HDF5, MPI, BLAS/LAPACK or CUDA (CMakeLists.txt's gwcore + gw targets).
REF_SOURCES = [
    "src/main.cpp", "src/cli.cpp", "src/execution.cpp", "src/io.cpp", "src/memory_footprint.cpp",
    "src/mpi_context.cpp", "src/profiling.cpp", "src/gw/frequency_grids.cpp", "src/gw/gw.cpp",
    "src/gw/pade.cpp", "src/gw/qp_solver.cpp", "src/gw/sigma.cpp", "src/gw/sigma_c_driver.cpp",
    "src/linalg/linalg.cpp", "src/linalg/backend_factory.cpp", "src/runtime/kernel_policy.cpp",
    "src/workspace/screening_workspace.cpp", "src/workspace/pq_ph_panel.cpp",
]
CXXFLAGS = ["-std=c++20", "-O1", "-fopenmp", "-DNDEBUG", "-DGW_ENABLE_OPENMP_KERNEL_LOOPS",
            "-DGW_HAS_HIP_BACKEND"]


def guide_blocks(path=GUIDE):
    """(kind, name, text) of every cpp file= block and every diff block."""
    text = open(path).read()
    out = [("cpp", n, body) for n, body in re.findall(r"```cpp file=(\S+)\n(.*?)```", text, re.S)]
    for body in re.findall(r"```diff\n(.*?)```", text, re.S):
        m = re.search(r"^\+\+\+ b/(\S+)", body, re.M)
        if not m:
            raise ValueError("diff block without a '+++ b/<file>' header")
        out.append(("diff", m.group(1), body))
    return out


def prepare_tree(ref, work):
    """Writes the guide's files and the patched reference files under work;
    returns the list of .cpp files to compile."""
    patched = set()
    sources = []
    for kind, name, body in guide_blocks():
        if kind == "cpp":
            dst = os.path.join(work, name)
            os.makedirs(os.path.dirname(dst), exist_ok=True)
            with open(dst, "w") as f:
                f.write(body)
            if name.endswith(".cpp"):
                sources.append(dst)
        else:
            dst = os.path.join(work, name)
            if not os.path.exists(dst):
                os.makedirs(os.path.dirname(dst), exist_ok=True)
                shutil.copyfile(os.path.join(ref, name), dst)
            res = subprocess.run(["patch", "-p1", "--batch", "--forward", "-d", work], input=body,
                                 capture_output=True, text=True)
            if res.returncode != 0:
                raise RuntimeError(f"patch of {name} failed:\n{res.stdout}{res.stderr}")
            patched.add(name)
    for name in REF_SOURCES:
        sources.append(os.path.join(work, name) if name in patched else os.path.join(ref, name))
    return sources


def build(out=DEFAULT_OUT, ref="/root/reference", jobs=8, cxx="g++", log=print):
    lib = os.path.join(LIB_DIR, "libmgw_hip.so")
    if not os.path.exists(lib):
        raise FileNotFoundError(f"{lib} is not built (python minigw_amd/build.py)")
    work = tempfile.mkdtemp(prefix="mgw_integration_")
    try:
        sources = prepare_tree(ref, work)
        incs = ["-I" + os.path.join(work, "include"), "-I" + os.path.join(ref, "include"),
                "-I" + os.path.join(REPO, "include")]

        def compile_one(i_src):
            i, src = i_src
            obj = os.path.join(work, f"obj{i:03d}.o")
            res = subprocess.run([cxx] + CXXFLAGS + incs + ["-c", src, "-o", obj],
                                 capture_output=True, text=True)
            if res.returncode != 0:
                raise RuntimeError(f"compiling {src} failed:\n{res.stderr}")
            return obj

        with ThreadPoolExecutor(max_workers=max(1, jobs)) as pool:
            objs = list(pool.map(compile_one, enumerate(sources)))
        os.makedirs(os.path.dirname(os.path.abspath(out)), exist_ok=True)
        rpath = os.path.relpath(LIB_DIR, os.path.dirname(os.path.abspath(out)))
        cmd = [cxx, "-fopenmp"] + objs + ["-o", out, "-L" + LIB_DIR, "-lmgw_hip",
                                          "-Wl,-rpath,$ORIGIN/" + rpath,
                                          "-Wl,-rpath-link,/opt/rocm/lib"]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"linking {out} failed:\n{res.stderr}")
        log(f"built {out} from {len(sources)} translation units")
        return out
    finally:
        shutil.rmtree(work, ignore_errors=True)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--ref", default="/root/reference")
    ap.add_argument("-j", type=int, default=8)
    a = ap.parse_args(argv)
    build(a.out, a.ref, a.j)
    return 0


if __name__ == "__main__":
    sys.exit(main())
