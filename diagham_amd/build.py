"""Build the native libraries in-tree.

    python -m diagham_amd.build            # libdhgpu.so + libdiagham_host.so (+ oracle)

libdhgpu.so       HIP kernels + the C ABI of include/dhgpu.h, compiled for gfx950
libdiagham_host.so C++ host mirroring the DiagHam classes on that path (include/diagham_host.h)
"""
import concurrent.futures
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "diagham_amd", "csrc")
HOST = os.path.join(ROOT, "diagham_amd", "host")
OBJ = os.path.join(ROOT, "diagham_amd", "_build")
LIB_GPU = os.path.join(ROOT, "diagham_amd", "libdhgpu.so")
LIB_HOST = os.path.join(ROOT, "diagham_amd", "libdiagham_host.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")  # the host layer has no device code
ARCH = os.environ.get("DH_OFFLOAD_ARCH", "gfx950")
# packed FP32 instructions are disabled for every device compile: some of their forms have given
# wrong results on this hardware without any error, so the compiler must not emit them
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH, "-munsafe-fp-atomics",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-ffp-contract=fast", "-Wno-unused-result"]
GPU_SOURCES = ["dh_core.hip", "dh_basis.hip", "dh_hv.hip", "dh_aux.hip", "dh_multisector.hip", "dh_lattice.hip",
               "dh_nbody.hip", "dh_torus.hip", "dh_update.hip", "dh_noalloc.hip"]
HOST_SOURCES = ["clebsch_gordan.cc", "fermion_on_sphere.cc", "hamiltonians.cc", "lanczos.cc", "sectors.cc",
                "lattice.cc", "torus.cc", "vector_io.cc", "host_api.cc"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed: %s\n%s" % (" ".join(cmd), r.stdout))
    return r.stdout


def _compile_hip(src):
    out = os.path.join(OBJ, os.path.basename(src) + ".o")
    _run([HIPCC] + HIP_FLAGS + ["-c", src, "-o", out])
    return out


def _compile_host(src):
    out = os.path.join(OBJ, os.path.basename(src) + ".o")
    _run([CXX, "-O3", "-std=c++17", "-fPIC", "-fopenmp", "-Wall", "-Wno-unused-function",
          "-I" + os.path.join(ROOT, "include"), "-c", src, "-o", out])
    return out


def build_gpu(jobs=8):
    os.makedirs(OBJ, exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in GPU_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile_hip, srcs))
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", LIB_GPU + ".tmp"] + objs)
    os.replace(LIB_GPU + ".tmp", LIB_GPU)
    return LIB_GPU


def build_host(jobs=8):
    os.makedirs(OBJ, exist_ok=True)
    srcs = [os.path.join(HOST, s) for s in HOST_SOURCES if os.path.exists(os.path.join(HOST, s))]
    if not srcs:
        return None
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile_host, srcs))
    _run([CXX, "-shared", "-fPIC", "-fopenmp", "-o", LIB_HOST + ".tmp"] + objs +
         ["-L" + os.path.dirname(LIB_GPU), "-ldhgpu", "-Wl,-rpath,$ORIGIN", "-lquadmath"])
    os.replace(LIB_HOST + ".tmp", LIB_HOST)
    return LIB_HOST


def build_oracle():
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])


def build_all(jobs=8):
    build_gpu(jobs)
    build_host(jobs)
    build_oracle()


if __name__ == "__main__":
    build_all(int(sys.argv[1]) if len(sys.argv) > 1 else 8)
    print("built", LIB_GPU, LIB_HOST)
