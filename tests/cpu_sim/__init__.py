"""CPU build of the whole libgst (test infrastructure).

``gpugst_amd/csrc/gst_device.hip`` and ``gst_host.cpp`` are compiled with g++
against ``hip/hip_runtime.h`` (every lane of a wave is a host thread; wave and
block operations synchronise those threads) and ``mock_runtime.cpp`` (host
memory stands in for device memory; every copy and launch completes when it
is issued).  The result exports the same C ABI as the GPU build, so the
ctypes binding loads it through the ``GST_LIB`` environment variable.  CPU
tests use it to run the engine's device logic and host engine end to end
where no GPU exists; it is never a product path.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(ROOT, "gpugst_amd", "csrc")
LIB = os.path.join(HERE, "_build", "libgst_cpu.so")
SOURCES = [os.path.join(HERE, "cpu_lib.cpp"), os.path.join(CSRC, "gst_host.cpp"),
           os.path.join(HERE, "mock_runtime.cpp")]
DEPS = SOURCES + [os.path.join(CSRC, "gst_device.hip"), os.path.join(CSRC, "gst_common.h"),
                  os.path.join(HERE, "hip", "hip_runtime.h"), os.path.join(ROOT, "include", "gst.h")]


def build(flags=("-O1", "-g")) -> str:
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    tmp = LIB + ".tmp"
    cmd = ["g++", "-std=c++20", *flags, "-fPIC", "-shared", "-pthread", "-DGST_CPU_SIM",
           "-DGST_CPU_SIM_HOST", "-I", HERE, "-I", CSRC, "-I", os.path.join(ROOT, "include"),
           "-x", "c++", *SOURCES, "-o", tmp]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


def lib_path() -> str:
    """Path of an up-to-date CPU build (rebuilt when a source is newer)."""
    if not os.path.exists(LIB) or any(os.path.getmtime(LIB) < os.path.getmtime(d) for d in DEPS):
        build()
    return LIB
