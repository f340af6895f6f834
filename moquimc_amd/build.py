"""Build libmqi_hip.so for gfx950 in-tree (moquimc_amd/libmqi_hip.so).

    python -m moquimc_amd.build            # or: python moquimc_amd/build.py

Translation units:
  mqi_transport.hip  -> twice: MQI_EXACT=1 (-ffp-contract=off, portable math;
                        the verification kernels) and MQI_EXACT=0 (production)
  mqi_aux.hip        -> fills, calibration, statistics, Dij CSR, sparse products
  mqi_api.cpp        -> the C ABI of include/mqi_hip.h
"""
import os
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libmqi_hip.so")
ARCH = os.environ.get("MQI_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC, "-Wno-unused-result", "-Wno-unused-value",
]


def hipcc():
    p = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    return p


def _compile(args):
    src, obj, extra = args
    cmd = [hipcc(), "-c"] + COMMON + extra + [src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


def build(verbose=True):
    tmp = tempfile.mkdtemp(prefix="mqi_build_")
    try:
        jobs = [
            (os.path.join(CSRC, "mqi_transport.hip"), os.path.join(tmp, "transport_exact.o"),
             ["-DMQI_EXACT=1", "-ffp-contract=off"]),
            (os.path.join(CSRC, "mqi_transport.hip"), os.path.join(tmp, "transport_fast.o"),
             ["-DMQI_EXACT=0", "-ffp-contract=fast-honor-pragmas",
              "-fno-hip-fp32-correctly-rounded-divide-sqrt"]),
            (os.path.join(CSRC, "mqi_aux.hip"), os.path.join(tmp, "aux.o"), ["-ffp-contract=off"]),
            (os.path.join(CSRC, "mqi_api.cpp"), os.path.join(tmp, "api.o"), ["-x", "hip", "-ffp-contract=off"]),
        ]
        with ThreadPoolExecutor(max_workers=4) as ex:
            objs = list(ex.map(_compile, jobs))
        out_tmp = LIB + ".tmp"
        cmd = [hipcc(), "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", out_tmp] + objs
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("link failed: %s\n%s" % (r.stdout, r.stderr))
        os.replace(out_tmp, LIB)
        if verbose:
            print("built", LIB)
        return LIB
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def build_host(verbose=True):
    """The C++ phantom driver over the library (moquimc_amd/host)."""
    host = os.path.join(os.path.dirname(LIB), "host")
    exe = os.path.join(host, "phantom_env")
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-I" + os.path.join(ROOT, "include"), "-I" + host,
           os.path.join(host, "phantom_main.cpp"), "-L" + os.path.dirname(LIB), "-lmqi_hip",
           "-Wl,-rpath,$ORIGIN/..", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("host build failed: %s\n%s" % (" ".join(cmd), r.stderr))
    if verbose:
        print("built", exe)
    return exe


if __name__ == "__main__":
    build()
    build_host()
    sys.exit(0)
