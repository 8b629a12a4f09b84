"""Build libcatana_hip.so (HIP, gfx950) and the pybind11 module in-tree.

    python catana_amd/build.py          # everything
"""
import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# gfx950 only; packed FP32 instructions are switched off for all device code.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
    "-Wall", "-Wno-unused-function",
]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_lib():
    out = os.path.join(HERE, "libcatana_hip.so")
    _run([HIPCC, *HIP_FLAGS, "-shared", os.path.join(CSRC, "sfb_capi.hip"), "-o", out])
    return out


def build_pymodule():
    import pybind11
    ext = sysconfig.get_config_var("EXT_SUFFIX")
    out = os.path.join(HERE, "_catana" + ext)
    inc = [pybind11.get_include(), sysconfig.get_paths()["include"]]
    _run([HIPCC, "-O2", "-std=c++17", "-fPIC", "-shared", *["-I" + i for i in inc],
          "-I" + os.path.join(ROOT, "include"), "-I" + CSRC, "-D__HIP_PLATFORM_AMD__",
          os.path.join(CSRC, "catana_api.cpp"), os.path.join(CSRC, "py_catana.cpp"),
          "-L" + HERE, "-lcatana_hip", "-Wl,-rpath,$ORIGIN", "-o", out])
    return out


def build_all():
    build_lib()
    if os.path.exists(os.path.join(CSRC, "py_catana.cpp")):
        build_pymodule()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
