"""In-tree build of the native libraries.

* ``gpugst_amd/_lib/libgst.so``     -- the engine: HIP kernels for gfx950 plus
  the host engine and the C ABI of ``include/gst.h`` (hipcc).
* ``gpugst_amd/_lib/libgstdata.so`` -- host-only synthetic-data generators,
  CSR construction and the readers of the reference's text formats (g++).

Run ``python -m gpugst_amd.build``.  Nothing here needs a GPU.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_lib")
LIBGST = os.path.join(OUT, "libgst.so")
LIBDATA = os.path.join(OUT, "libgstdata.so")

ARCH = os.environ.get("GST_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", shutil.which("g++") or "g++")

# Packed-FP32 instructions are disabled for all device code (the engine has
# no floating point on its hot path anyway).
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-function"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=HERE, capture_output=True, text=True)
    # hipcc's host pass reports the device-only target feature; that is noise
    err = "\n".join(line for line in r.stderr.splitlines()
                    if "packed-fp32-ops" not in line)
    if err.strip():
        print(err, file=sys.stderr)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}")


def build(verbose: bool = True) -> None:
    os.makedirs(OUT, exist_ok=True)
    inc = ["-I", os.path.join(ROOT, "include"), "-I", CSRC]
    objs = []
    for src in ("gst_device.hip", "gst_host.cpp"):
        obj = os.path.join(OUT, src.rsplit(".", 1)[0] + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        _run([HIPCC, *HIP_FLAGS, *inc, *lang, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIBGST + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", *objs, "-o", tmp,
          "-lpthread"])
    os.replace(tmp, LIBGST)
    tmp = LIBDATA + ".tmp"
    _run([CXX, "-O3", "-std=c++17", "-fPIC", "-shared", "-Wall",
          os.path.join(CSRC, "gst_data.cpp"), os.path.join(CSRC, "gst_formats.cpp"), "-o", tmp,
          "-lpthread"])
    os.replace(tmp, LIBDATA)
    for o in objs:
        os.remove(o)


if __name__ == "__main__":
    build()
