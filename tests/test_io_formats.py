"""Output writers (moquimc_amd/host/mqi_hip_io.hpp) in the reference's
formats (moqui/base/mqi_io.hpp:165-183,249-320,493-591 and the zip writer
mqi_sparse_io.hpp:166-400), read back by the readers users have: scipy's
load_npz, zipfile (which checks every member's CRC-32), numpy.load, and a
MetaImage header parser.  The CRC-32 is pinned by its published check value
(crc32("123456789") = 0xCBF43926) and by zlib's on random bytes."""
import binascii
import os
import subprocess
import zipfile

import numpy as np
import pytest
import scipy.sparse as sps

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("io") / "io_driver")
    subprocess.run(["g++", "-std=c++17", "-O2", "-I", os.path.join(ROOT, "moquimc_amd", "host"),
                    os.path.join(ROOT, "tests", "io_driver.cpp"), "-o", exe], check=True)
    return exe


def _run(*args):
    subprocess.run([str(a) for a in args], check=True)


def test_crc32_check_value_and_zlib(driver, tmp_path):
    f = tmp_path / "check.bin"
    f.write_bytes(b"123456789")
    out = subprocess.run([driver, "crc", str(f)], capture_output=True, text=True, check=True).stdout.strip()
    assert out == "cbf43926"
    data = np.random.default_rng(0).integers(0, 256, 100003, dtype=np.uint8).tobytes()
    f.write_bytes(data)
    out = subprocess.run([driver, "crc", str(f)], capture_output=True, text=True, check=True).stdout.strip()
    assert int(out, 16) == binascii.crc32(data)


def _csr(rng, rows, cols, density=0.05):
    m = sps.random(rows, cols, density=density, format="csr", random_state=np.random.RandomState(1),
                   dtype=np.float32)
    m.data = (m.data + 0.1).astype(np.float32)
    m.sort_indices()
    return m


def test_npz_reads_back_with_scipy(driver, tmp_path):
    m = _csr(np.random.default_rng(1), 37, 513)
    m.indptr = m.indptr.astype(np.int64)
    ip, ix, v = tmp_path / "ip.bin", tmp_path / "ix.bin", tmp_path / "v.bin"
    m.indptr.astype(np.int64).tofile(ip)
    m.indices.astype(np.int32).tofile(ix)
    m.data.astype(np.float32).tofile(v)
    out = tmp_path / "Dij.npz"
    scale = 2.5e7
    _run(driver, "npz", out, 37, 513, repr(scale), ip, ix, v)
    with zipfile.ZipFile(out) as z:
        assert z.testzip() is None                   # every member's CRC-32 checks
        assert z.namelist() == ["indices.npy", "indptr.npy", "shape.npy", "data.npy", "format.npy"]
        assert all(i.compress_type == zipfile.ZIP_STORED for i in z.infolist())
    with np.load(out) as f:
        assert f["indices"].dtype == np.uint32 and f["indptr"].dtype == np.uint32
        assert f["data"].dtype == np.float64 and f["shape"].dtype == np.uint32
        assert list(f["shape"]) == [37, 513]
        assert f["format"].item() == b"csr"
    got = sps.load_npz(out)
    assert got.shape == (37, 513) and got.format == "csr"
    assert np.array_equal(got.indptr, m.indptr) and np.array_equal(got.indices, m.indices)
    assert np.array_equal(got.data, m.data.astype(np.float64) * scale)


def test_npz_empty_rows_and_empty_matrix(driver, tmp_path):
    for rows, cols, dens in ((5, 40, 0.0), (6, 50, 0.2)):
        m = _csr(None, rows, cols, dens)
        if dens:
            m = sps.csr_matrix(m.toarray() * (np.arange(rows)[:, None] % 2))   # odd rows only
        ip, ix, v = tmp_path / "ip.bin", tmp_path / "ix.bin", tmp_path / "v.bin"
        m.indptr.astype(np.int64).tofile(ip)
        m.indices.astype(np.int32).tofile(ix)
        m.data.astype(np.float32).tofile(v)
        out = tmp_path / ("m%d.npz" % rows)
        _run(driver, "npz", out, rows, cols, "1.0", ip, ix, v)
        got = sps.load_npz(out)
        assert got.shape == (rows, cols) and got.nnz == m.nnz
        assert np.array_equal(got.toarray(), m.toarray().astype(np.float64))


def _meta(path):
    """key -> value of a MetaImage header (both 'Key = v' and 'Key v'), and
    the byte offset where the header ends."""
    raw = open(path, "rb").read()
    keys, pos = {}, 0
    while True:
        end = raw.index(b"\n", pos)
        line = raw[pos:end].decode()
        pos = end + 1
        k, _, v = line.partition(" = ") if " = " in line else line.partition(" ")
        keys[k] = v
        if k == "ElementDataFile":
            return keys, pos


def test_dense_raw_mhd_mha(driver, tmp_path):
    nx, ny, nz = 7, 5, 3
    xe = np.linspace(-3.5, 3.5, nx + 1, dtype=np.float32) * np.float32(0.98)
    ye = np.linspace(-10.0, 0.0, ny + 1, dtype=np.float32)
    ze = np.linspace(0.0, 6.0, nz + 1, dtype=np.float32)
    d = np.random.default_rng(2).random(nx * ny * nz)
    files = []
    for name, a in (("v", d), ("xe", xe), ("ye", ye), ("ze", ze)):
        p = tmp_path / (name + ".bin")
        a.tofile(p)
        files.append(p)
    scale = 3.0
    _run(driver, "dense", tmp_path, "Dose", nx, ny, nz, repr(scale), *files)
    want = d * scale
    assert np.array_equal(np.fromfile(tmp_path / "Dose_raw.raw", np.float64), want)
    keys, _ = _meta(tmp_path / "Dose_mhd.mhd")
    assert keys["ElementDataFile"] == "Dose_mhd.raw" and keys["ElementType"] == "MET_DOUBLE"
    assert keys["DimSize"].split() == ["7", "5", "3"]
    assert keys["TransformMatrix"] == "1 0 0 0 1 0 0 0 1" and keys["BinaryDataByteOrderMSB"] == "False"
    # spacing = e[1] - e[0], offset = first voxel centre (6 significant digits in .mhd)
    sp = [float(x) for x in keys["ElementSpacing"].split()]
    off = [float(x) for x in keys["Offset"].split()]
    for e, s, o in zip((xe, ye, ze), sp, off):
        assert abs(s - (e[1] - e[0])) <= 1e-5 * abs(e[1] - e[0])
        assert abs(o - (e[0] + 0.5 * (e[1] - e[0]))) <= 1e-5 * max(1.0, abs(o))
    assert np.array_equal(np.fromfile(tmp_path / "Dose_mhd.raw", np.float64), want)
    keys, pos = _meta(tmp_path / "Dose_mha.mha")
    assert keys["ElementDataFile"] == "LOCAL" and keys["HeaderSize"] == "-1"
    org = [float(x) for x in keys["Origin"].split()]
    assert np.allclose(org, off, rtol=1e-5, atol=1e-5)
    data = np.frombuffer(open(tmp_path / "Dose_mha.mha", "rb").read()[pos:], np.float64)
    assert np.array_equal(data, want)


def _write_ct(path, hu, origin, spacing, element_type="MET_FLOAT", key_case=str):
    nz, ny, nx = hu.shape
    hdr = ["ObjectType = Image", "NDims = 3", "BinaryData = True", "BinaryDataByteOrderMSB = False",
           "CompressedData = False", "TransformMatrix = 1 0 0 0 1 0 0 0 1",
           key_case("Offset") + " = %r %r %r" % tuple(origin), "CenterOfRotation = 0 0 0",
           key_case("ElementSpacing") + " = %r %r %r" % tuple(spacing), "DimSize = %d %d %d" % (nx, ny, nz),
           "ElementType = " + element_type, "ElementDataFile = LOCAL"]
    with open(path, "wb") as f:
        f.write(("\n".join(hdr) + "\n").encode())
        f.write(hu.astype(np.float32).tobytes())


def _edges(o, d, n, shift=0.0):
    # x0 = origin - d / 2.0 in double, rounded to float; + shift in float;
    # x_i = x0 + i * d in float
    f32 = np.float32
    x0 = f32(f32(o) - f32(d) / 2.0)
    x0 = f32(x0 + f32(shift))
    return np.array([f32(x0 + f32(f32(i) * f32(d))) for i in range(n + 1)], np.float32)


def test_mha_ct_reader(driver, tmp_path):
    rng = np.random.default_rng(3)
    hu = rng.uniform(-1000, 2000, (4, 5, 6)).astype(np.float32)   # non-integer values: truncation
    origin, spacing = (-10.25, 3.5, -100.0), (0.98, 1.25, 2.0)
    p = tmp_path / "ct.mha"
    _write_ct(p, hu, origin, spacing, key_case=str.upper)
    out = subprocess.run([driver, "ct", str(p), str(tmp_path / "hu.bin")], capture_output=True, text=True,
                         check=True).stdout.splitlines()
    assert out[0].split() == ["6", "5", "4"]
    xe, ye, ze = ([float.fromhex(v) for v in ln.split()] for ln in out[1:4])
    assert np.array_equal(np.float32(xe), _edges(origin[0], spacing[0], 6))
    assert np.array_equal(np.float32(ye), _edges(origin[1], spacing[1], 5))
    # z accumulates: z_i = z_{i-1} + dz
    z = [np.float32(np.float32(origin[2]) - np.float32(spacing[2]) / 2.0)]
    for _ in range(4):
        z.append(np.float32(z[-1] + np.float32(spacing[2])))
    assert np.array_equal(np.float32(ze), np.float32(z))
    # the first voxel centre is the Offset, to rounding
    assert abs(0.5 * (xe[0] + xe[1]) - origin[0]) < 1e-5
    got = np.fromfile(tmp_path / "hu.bin", np.int16)
    assert np.array_equal(got, np.trunc(hu).astype(np.int16).ravel())
    # a volume shift moves the first edge
    out = subprocess.run([driver, "ct", str(p), str(tmp_path / "hu.bin"), "1.5", "-2", "0"], capture_output=True,
                         text=True, check=True).stdout.splitlines()
    assert np.array_equal(np.float32([float.fromhex(v) for v in out[1].split()]), _edges(origin[0], spacing[0], 6, 1.5))


def test_mha_ct_reader_refuses_bad_files(driver, tmp_path):
    hu = np.zeros((2, 2, 2), np.float32)
    p = tmp_path / "bad.mha"
    _write_ct(p, hu, (0, 0, 0), (1, 1, 1), element_type="MET_SHORT")
    assert subprocess.run([driver, "ct", str(p), str(tmp_path / "o.bin")], capture_output=True).returncode == 1
    _write_ct(p, hu, (0, 0, 0), (1, 1, 1))
    data = open(p, "rb").read()
    open(p, "wb").write(data[:-4])   # one voxel short
    assert subprocess.run([driver, "ct", str(p), str(tmp_path / "o.bin")], capture_output=True).returncode == 1


def write_mask(path, m, element_type="MET_UCHAR", local=True):
    """A structure mask as MetaImage (uint8 after the header), shape (nz, ny, nx)."""
    nz, ny, nx = m.shape
    hdr = ["ObjectType = Image", "NDims = 3", "BinaryData = True", "BinaryDataByteOrderMSB = False",
           "CompressedData = False", "TransformMatrix = 1 0 0 0 1 0 0 0 1", "Offset = 0 0 0",
           "CenterOfRotation = 0 0 0", "AnatomicalOrientation = RAI", "ElementSpacing = 1 1 1",
           "DimSize = %d %d %d" % (nx, ny, nz), "ElementType = " + element_type,
           "ElementDataFile = " + ("LOCAL" if local else "mask.raw")]
    with open(path, "wb") as f:
        f.write(("\n".join(hdr) + "\n").encode())
        f.write(np.ascontiguousarray(m, np.uint8).tobytes())


def test_mask_files_sum_to_mask_total(driver, tmp_path):
    """mask_reader::read_mask_files: the structure masks summed voxel by
    voxel (overlaps give 2); refused: another grid size, a value above 1,
    a wider element type, data outside the file."""
    rng = np.random.default_rng(4)
    shape = (5, 6, 7)
    a = (rng.random(shape) < 0.3).astype(np.uint8)
    b = (rng.random(shape) < 0.4).astype(np.uint8)
    pa, pb, out = tmp_path / "a.mha", tmp_path / "b.mha", tmp_path / "total.u8"
    write_mask(pa, a)
    write_mask(pb, b)
    _run(driver, "mask", out, 7, 6, 5, pa, pb)
    got = np.fromfile(out, np.uint8)
    assert np.array_equal(got, (a.astype(int) + b).ravel())
    assert got.max() == 2
    bad = tmp_path / "bad.mha"
    for m, kw, dims in ((a, {}, (7, 6, 4)), (a * 2, {}, (7, 6, 5)), (a, {"element_type": "MET_SHORT"}, (7, 6, 5)),
                        (a, {"local": False}, (7, 6, 5))):
        write_mask(bad, m, **kw)
        r = subprocess.run([driver, "mask", str(out)] + [str(d) for d in dims] + [str(bad)], capture_output=True)
        assert r.returncode == 1, (kw, dims)
