"""The C ABI on a machine without a GPU: the library loads, exports every
function include/mqi_hip.h declares, the ctypes mirrors of the ABI structs
have the C layout (checked against a program compiled from the header), and
a context on a missing device fails with an error code and a message instead
of exiting.
"""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mqi_hip.h")
LIB = os.path.join(ROOT, "moquimc_amd", "libmqi_hip.so")

needs_lib = pytest.mark.skipif(not os.path.exists(LIB), reason="libmqi_hip.so not built (python -m moquimc_amd.build)")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mqi_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_abi():
    names = _declared()
    for must in ("mqi_create", "mqi_destroy", "mqi_last_error", "mqi_run", "mqi_dij_matvec"):
        assert must in names
    from moquimc_amd import _abi
    assert sorted(_abi.EXPORTED) == names


@needs_lib
def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    missing = [n for n in _declared() if n not in exported]
    assert not missing, missing
    lib = ctypes.CDLL(LIB)
    for n in _declared():
        assert getattr(lib, n) is not None


def _c_layout():
    """sizeof / offsetof of the ABI structs from a host program compiled
    against the header with gcc."""
    src = r"""
#include <stdio.h>
#include <stddef.h>
#include "mqi_hip.h"
#define F(S, m) printf(#S "." #m " %zu\n", offsetof(S, m));
int main(void) {
    printf("mqi_spot %zu\nmqi_run_config %zu\nmqi_run_result %zu\n", sizeof(mqi_spot), sizeof(mqi_run_config),
           sizeof(mqi_run_result));
    F(mqi_spot, mean) F(mqi_spot, sigma) F(mqi_spot, corr) F(mqi_spot, rotation) F(mqi_spot, translation)
    F(mqi_spot, histories)
    F(mqi_run_config, stat_threshold) F(mqi_run_config, max_rounds) F(mqi_run_config, dij)
    F(mqi_run_config, dij_capacity) F(mqi_run_config, hist_offset) F(mqi_run_config, dij_slots_sqrt)
    F(mqi_run_config, use_graph)
    printf("mqi_grid_desc %zu\n", sizeof(mqi_grid_desc));
    F(mqi_grid_desc, xe) F(mqi_grid_desc, density) F(mqi_grid_desc, rotation) F(mqi_grid_desc, translation)
    F(mqi_run_result, dropped) F(mqi_run_result, steps) F(mqi_run_result, status) F(mqi_run_result, kernel_ms)
    F(mqi_run_result, stat)
    return 0;
}
"""
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "layout.c")
        exe = os.path.join(d, "layout")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
    return {k: int(v) for k, v in (ln.split() for ln in out.splitlines())}


def test_ctypes_mirrors_match_c_layout():
    from moquimc_amd import _abi
    lay = _c_layout()
    assert lay["mqi_spot"] == _abi.SPOT_DTYPE.itemsize == 128
    for f in ("mean", "sigma", "corr", "rotation", "translation", "histories"):
        assert lay["mqi_spot." + f] == _abi.SPOT_DTYPE.fields[f][1], f
    assert lay["mqi_run_config"] == ctypes.sizeof(_abi.RunConfig)
    assert lay["mqi_run_result"] == ctypes.sizeof(_abi.RunResult)
    for f in ("stat_threshold", "max_rounds", "dij", "dij_capacity", "hist_offset", "dij_slots_sqrt", "use_graph"):
        assert lay["mqi_run_config." + f] == getattr(_abi.RunConfig, f).offset, f
    assert lay["mqi_grid_desc"] == ctypes.sizeof(_abi.GridDesc)
    for f in ("xe", "density", "rotation", "translation"):
        assert lay["mqi_grid_desc." + f] == getattr(_abi.GridDesc, f).offset, f
    for f in ("dropped", "steps", "status", "kernel_ms", "stat"):
        assert lay["mqi_run_result." + f] == getattr(_abi.RunResult, f).offset, f


@needs_lib
def test_create_without_device_reports_error():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    lib = ctypes.CDLL(LIB)
    lib.mqi_last_error.restype = ctypes.c_char_p
    lib.mqi_last_error.argtypes = [ctypes.c_void_p]
    h = ctypes.c_void_p()
    rc = lib.mqi_create(0, ctypes.byref(h))
    assert rc != 0
    assert not h.value
    msg = lib.mqi_last_error(None)
    assert msg and len(msg) > 0


def test_null_context_calls_fail_cleanly():
    if not os.path.exists(LIB):
        pytest.skip("libmqi_hip.so not built")
    lib = ctypes.CDLL(LIB)
    assert lib.mqi_destroy(None) in (0, 1)
    assert lib.mqi_zero_scorers(None) == 1
    assert lib.mqi_run(None, None, None) == 1


HOST = os.path.join(ROOT, "moquimc_amd", "host")


def test_host_layer_rotation_matches_oracle():
    """mqi_hip::coordinate_rotation (C++ environment layer) equals the
    oracle's restatement of coordinate_transform, which equals the
    reference's (spot sampling tests), bit for bit."""
    import tempfile
    import refbind as rb
    angles = [(0, 0, 0, 0), (0, 0, 0, 90), (15, 30, 10, 90), (-20, 45.5, 270, 90), (90, 180, 0, 0)]
    src = "#include <cstdio>\n#include \"mqi_hip_env.hpp\"\nint main(){float a[4];\n" \
          "while (std::scanf(\"%f %f %f %f\", a, a+1, a+2, a+3) == 4) {\n" \
          "auto r = mqi_hip::coordinate_rotation({a[0], a[1], a[2], a[3]});\n" \
          "for (float v : r) std::printf(\"%a \", v); std::printf(\"\\n\"); } return 0; }\n"
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "rot.cpp")
        exe = os.path.join(d, "rot")
        open(c, "w").write(src)
        subprocess.run(["g++", "-std=c++17", "-O2", "-I", os.path.join(ROOT, "include"), "-I", HOST, c, "-o", exe],
                       check=True)
        inp = "\n".join(" ".join(str(x) for x in a) for a in angles) + "\n"
        out = subprocess.run([exe], input=inp, capture_output=True, text=True, check=True).stdout.split("\n")
    for a, line in zip(angles, out):
        got = np.array([float.fromhex(t) for t in line.split()], np.float32)
        want = rb.coordinate_rotation(a)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), a


@needs_lib
def test_host_driver_fails_cleanly_without_device():
    import torch
    exe = os.path.join(HOST, "phantom_env")
    if not os.path.exists(exe):
        pytest.skip("phantom_env not built (python -m moquimc_amd.build)")
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    r = subprocess.run([exe, "--output_prefix", "/nonexistent/", "--nxyz", "4", "4", "4"], capture_output=True,
                       text=True)
    assert r.returncode == 1
    assert "mqi_create" in r.stderr
    r = subprocess.run([exe, "--nxyz", "4", "4"], capture_output=True, text=True)
    assert r.returncode == 1 and "--nxyz needs 3 values" in r.stderr


@needs_lib
def test_ct_env_driver_builds_and_fails_cleanly_without_device(tmp_path):
    """The tps_env-style C++ environment (mqi_hip_ct_env.hpp) compiles and
    links against the library; without a device, initialize() reports
    mqi_create's error and the driver exits with status 1."""
    import torch
    exe = str(tmp_path / "ct_env_driver")
    lib = os.path.join(ROOT, "moquimc_amd")
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), "-I", HOST,
                    os.path.join(ROOT, "tests", "ct_env_driver.cpp"), "-L", lib, "-lmqi_hip", "-Wl,-rpath," + lib,
                    "-o", exe], check=True)
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    spots = tmp_path / "spots.bin"
    spots.write_bytes(b"")
    r = subprocess.run([exe, "/nonexistent.mha", str(spots), str(tmp_path), "raw", "1", "Dose", "0", "-"],
                       capture_output=True, text=True)
    assert r.returncode == 1 and r.stderr.strip()


def test_ct_env_output_scale_is_the_float_product(tmp_path):
    """ct_env scales its outputs by particles_per_history * rbe * n_fractions
    evaluated as tps_env does, in float with an int fraction count
    (mqi_tps_env.hpp:117,133-134,1873-1889): equal to numpy's float32
    product, and for RBE 1.1 over 30 fractions different from the double
    product (so the test sees the difference)."""
    src = tmp_path / "scale.cpp"
    src.write_text('#include <cstdio>\n#include "mqi_hip_ct_env.hpp"\n'
                   'int main(){ const float p[] = {1.0f, 1.0f, 0.5f, 1e8f, 3.7e4f};\n'
                   ' const float r[] = {1.1f, 1.1f, 1.0f, 1.1f, 1.05f}; const int n[] = {1, 30, 7, 33, 19};\n'
                   ' for (int i = 0; i < 5; ++i) std::printf("%.17g\\n", mqi_hip::output_scale(p[i], r[i], n[i]));\n'
                   ' return 0; }\n')
    exe = str(tmp_path / "scale")
    lib = os.path.join(ROOT, "moquimc_amd")
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), "-I", HOST,
                    str(src), "-L", lib, "-lmqi_hip", "-Wl,-rpath," + lib, "-o", exe], check=True)
    got = [float(x) for x in subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()]
    p = np.float32([1.0, 1.0, 0.5, 1e8, 3.7e4])
    r = np.float32([1.1, 1.1, 1.0, 1.1, 1.05])
    n = np.float32([1, 30, 7, 33, 19])
    want = (p * r * n).astype(np.float64)
    assert got == list(want)
    assert got[1] == 33.0 and float(np.float32(1.1)) * 30 != 33.0   # the double product of the same inputs


def test_dij_size_guard_on_host(tmp_path):
    """mqi_dij_build refuses a scenario's Dij of more than 2^31 - 1 nonzeros
    with MQI_E_CAPACITY (its row sort counts items in 32-bit ints) before it
    allocates or sorts anything: the guard at the boundary, on the host."""
    src = tmp_path / "guard.cpp"
    src.write_text('#include <cstdio>\n#include "mqi_limits.h"\n'
                   'int main(){ const long long n[] = {0, 1, 2147483646ll, 2147483647ll, 2147483648ll, '
                   '4000000000ll, -1};\n for (long long v : n) std::printf("%d\\n", dij_size_check(v));\n'
                   ' return 0; }\n')
    exe = str(tmp_path / "guard")
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Werror", "-I", os.path.join(ROOT, "moquimc_amd", "csrc"),
                    str(src), "-o", exe], check=True)
    got = [int(x) for x in subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()]
    assert got == [0, 0, 0, 0, 5, 5, 5]   # MQI_OK ... MQI_E_CAPACITY


def test_csr_canonical_sorts_rows_and_sums_repeats():
    """A CSR in the reference's hash-table row order, with repeated columns
    and an empty row, becomes the canonical CSR scipy makes of it (columns
    ascending per row, repeats summed)."""
    import scipy.sparse as sps
    from moquimc_amd import _abi
    rng = np.random.default_rng(3)
    rows, cols = 40, 500
    lens = rng.integers(0, 30, rows)
    lens[7] = 0
    indptr = np.r_[0, np.cumsum(lens)].astype(np.uint32)
    indices = rng.integers(0, cols, int(indptr[-1])).astype(np.uint32)   # unsorted, with repeats
    data = rng.uniform(1e-6, 1.0, indices.size)
    ip, ix, v = _abi.csr_canonical(indptr, indices, data)
    want = sps.csr_matrix((data, indices.astype(np.int64), indptr.astype(np.int64)), shape=(rows, cols))
    want.sum_duplicates()
    want.sort_indices()
    assert np.array_equal(ip, want.indptr) and np.array_equal(ix, want.indices)
    assert np.array_equal(v, want.data.astype(np.float32))
    assert all(np.all(np.diff(ix[ip[r]:ip[r + 1]]) > 0) for r in range(rows))
