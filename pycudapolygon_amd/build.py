"""In-tree build of the native parts (gfx950):

  libpcp_hip.so           HIP kernels + C ABI (include/pcp.h)
  libpyCudaPolygon*.so    pybind11 module with the reference's Model API, linked to
                          libpcp_hip.so (rpath $ORIGIN)

Run `python -m pycudapolygon_amd.build` (or __graft_entry__.build()).  Device code
is compiled with -ffp-contract=off (no FMA contraction: every rounding happens as
written) and without packed-FP32 instructions.
"""
import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
ARCH = os.environ.get("PCP_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

LIB = os.path.join(HERE, "libpcp_hip.so")


def module_path():
    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    return os.path.join(HERE, "libpyCudaPolygon" + suffix)


DEVICE_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-but-set-variable",
]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build_lib():
    src = os.path.join(CSRC, "pcp_kernels.hip")
    tmp = LIB + ".tmp"
    _run([HIPCC, *DEVICE_FLAGS, "-shared", "-I", os.path.join(ROOT, "include"), "-o", tmp, src])
    os.replace(tmp, LIB)


def build_module():
    import pybind11
    out = module_path()
    tmp = out + ".tmp"
    inc_py = sysconfig.get_paths()["include"]
    srcs = [os.path.join(CSRC, "model.cpp"), os.path.join(CSRC, "pybind_module.cpp")]
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall",
           "-I", pybind11.get_include(), "-I", inc_py, "-I", os.path.join(ROOT, "include"),
           "-o", tmp, *srcs, "-L", HERE, "-lpcp_hip", "-Wl,-rpath,$ORIGIN"]
    _run(cmd)
    os.replace(tmp, out)


def build():
    build_lib()
    build_module()


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "all"
    if what in ("all", "lib"):
        build_lib()
    if what in ("all", "module"):
        build_module()
