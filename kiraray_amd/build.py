"""Build recipe of libkrr_hip.so (in-tree, gfx950).

The HIP translation unit is compiled with hipcc for gfx950 only, with
floating-point contraction off and correctly rounded fp32 division/sqrt
(the shading code must take the same decisions as its CPU checker) and
without packed fp32 instructions; the host-only sources are compiled with
g++.  The C++ render-pass layer (kiraray_amd/host/) becomes libkrr_host.so,
linked against libkrr_hip.so.  Outputs go to kiraray_amd/lib/.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB_DIR = os.path.join(HERE, "lib")
OBJ_DIR = os.path.join(LIB_DIR, "obj")
INCLUDES = ["-I" + os.path.join(ROOT, "include"), "-I" + CSRC]

HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
             "-fhip-fp32-correctly-rounded-divide-sqrt",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function"]

HIP_SOURCES = ["wavefront.hip", "nrc.hip", "passes.hip", "bvh_build.hip", "pathtracer.hip"]
CXX_SOURCES = ["scene.cpp", "bvh.cpp"]

# C++ render-pass layer over the C ABI (libkrr_host.so) and its test driver.
HOST_DIR = os.path.join(HERE, "host")
HOST_SOURCES = ["renderpass.cpp", "wavefront_pass.cpp", "megakernel_pass.cpp", "render_app.cpp", "passes.cpp"]
HOST_TEST = os.path.join(ROOT, "tests", "cpp", "host_tests.cpp")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HOST_FLAGS = CXX_FLAGS + ["-D__HIP_PLATFORM_AMD__", "-isystem", os.path.join(ROCM, "include"), "-I" + HOST_DIR]
HIP_RT = ["-L" + os.path.join(ROCM, "lib"), "-Wl,-rpath," + os.path.join(ROCM, "lib"), "-lamdhip64"]


def _hipcc():
    for c in ("/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found (ROCm is required to build kiraray_amd)")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "unknown target feature" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    if verbose and out.strip():
        print(out, flush=True)


def build(verbose=False):
    os.makedirs(OBJ_DIR, exist_ok=True)
    hipcc = _hipcc()
    jobs = []
    objs = []
    for s in HIP_SOURCES:
        o = os.path.join(OBJ_DIR, s.replace(".hip", ".o"))
        jobs.append([hipcc] + HIP_FLAGS + INCLUDES + ["-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    for s in CXX_SOURCES:
        o = os.path.join(OBJ_DIR, s.replace(".cpp", ".o"))
        jobs.append(["g++"] + CXX_FLAGS + INCLUDES + ["-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        list(ex.map(lambda c: _run(c, verbose), jobs))
    out = os.path.join(LIB_DIR, "libkrr_hip.so")
    tmp = out + ".tmp"
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC"] + objs + ["-o", tmp, "-lpthread"], verbose)
    os.replace(tmp, out)
    build_host(verbose)
    return out


def build_host(verbose=False):
    """libkrr_host.so (RenderPass / RenderApp layer) linked against
    libkrr_hip.so, and the tests/cpp driver linked against both."""
    os.makedirs(OBJ_DIR, exist_ok=True)
    objs = []
    jobs = []
    for s in HOST_SOURCES:
        o = os.path.join(OBJ_DIR, "host_" + s.replace(".cpp", ".o"))
        jobs.append(["g++"] + HOST_FLAGS + INCLUDES + ["-c", os.path.join(HOST_DIR, s), "-o", o])
        objs.append(o)
    with ThreadPoolExecutor(max_workers=len(jobs)) as ex:
        list(ex.map(lambda c: _run(c, verbose), jobs))
    link = ["-L" + LIB_DIR, "-Wl,-rpath,$ORIGIN", "-lkrr_hip"] + HIP_RT
    out = os.path.join(LIB_DIR, "libkrr_host.so")
    _run(["g++", "-shared", "-fPIC"] + objs + ["-o", out + ".tmp"] + link, verbose)
    os.replace(out + ".tmp", out)
    exe = os.path.join(LIB_DIR, "host_tests")
    _run(["g++"] + HOST_FLAGS + INCLUDES + [HOST_TEST, "-o", exe + ".tmp", "-L" + LIB_DIR, "-Wl,-rpath,$ORIGIN",
                                            "-lkrr_host", "-lkrr_hip"] + HIP_RT, verbose)
    os.replace(exe + ".tmp", exe)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
