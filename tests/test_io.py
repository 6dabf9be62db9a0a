"""On-disk formats (host code of the library; no device needed).

* VTK ImageData: the file is parsed back with an independent XML/base64
  reader and every point value is compared with the written arrays, in VTK's
  x-fastest point order.
* Checkpoint: a write/read round trip of populations and records, the
  header's descriptor fields, the per-cell record order of the reference's
  serializer (uPart, hydrodynamicForce, solidFraction, partId), and the
  rejection of a foreign or truncated file.
"""
import base64
import struct
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import lbdem_psm as L


def read_vti(path):
    root = ET.parse(path).getroot()
    assert root.get("type") == "ImageData" and root.get("header_type") == "UInt32"
    img = root.find("ImageData")
    ext = [int(v) for v in img.get("WholeExtent").split()]
    n = (ext[1] + 1, ext[3] + 1, ext[5] + 1)
    out = {}
    for da in img.find("Piece").find("PointData").findall("DataArray"):
        raw = base64.b64decode("".join(da.text.split()))
        nbytes = struct.unpack("<I", raw[:4])[0]
        vals = np.frombuffer(raw[4:4 + nbytes], dtype="<f4")
        m = int(da.get("NumberOfComponents"))
        # VTK point order: x fastest
        out[da.get("Name")] = vals.reshape(n[2], n[1], n[0], m)
    return n, float(img.get("Spacing").split()[0]), out


def test_vti_roundtrip(tmp_path):
    n = (5, 4, 3)
    nc = int(np.prod(n))
    rng = np.random.default_rng(0)
    rho = rng.normal(size=nc)
    u = rng.normal(size=(3, nc))
    p = tmp_path / "f.vti"
    L.vti_write(p, n, {"density": rho, "velocity": u}, spacing=0.25)
    nn, sp, got = read_vti(p)
    assert nn == n and sp == 0.25
    for x in range(n[0]):
        for y in range(n[1]):
            for z in range(n[2]):
                c = (x * n[1] + y) * n[2] + z
                assert got["density"][z, y, x, 0] == np.float32(rho[c])
                assert np.array_equal(got["velocity"][z, y, x], u[:, c].astype(np.float32))


def test_checkpoint_roundtrip_and_layout(tmp_path):
    n = (4, 3, 5)
    nc = int(np.prod(n))
    rng = np.random.default_rng(1)
    f = rng.normal(size=(19, nc))
    sf = rng.uniform(0, 1, nc)
    ids = rng.integers(0, 50, nc).astype(np.int64)
    us = rng.normal(size=(3, nc))
    F = rng.normal(size=(3, nc))
    d = L.make_desc(*n, periodic=(1, 0, 1), tau=0.73, precision="fp32_dev", slices_per_dim=8,
                    body_accel=(0, 0, 1e-6))
    p = tmp_path / "c.bin"
    L.checkpoint_write(p, d, f, sf, ids, us, F)
    d2, got = L.checkpoint_read(p)
    assert (d2.nx, d2.ny, d2.nz) == n and d2.tau == 0.73 and list(d2.periodic) == [1, 0, 1]
    assert d2.precision == L.FP32_DEV and d2.slices_per_dim == 8 and d2.body_accel[2] == 1e-6
    for k, a in (("f", f), ("sf", sf), ("id", ids), ("us", us), ("F", F)):
        assert np.array_equal(got[k], a), k
    # per-cell record: 19 populations, then uPart[3], force[3], sf, id
    raw = p.read_bytes()
    hdr = len(raw) - nc * 27 * 8
    cell = np.frombuffer(raw[hdr + 7 * 27 * 8: hdr + 8 * 27 * 8], dtype="<f8")
    assert np.array_equal(cell[:19], f[:, 7])
    assert np.array_equal(cell[19:22], us[:, 7]) and np.array_equal(cell[22:25], F[:, 7])
    assert cell[25] == sf[7] and np.frombuffer(cell[26:27].tobytes(), dtype="<i8")[0] == ids[7]
    # foreign and truncated files are refused
    (tmp_path / "bad.bin").write_bytes(b"x" * 200)
    with pytest.raises(L.PSMError):
        L.checkpoint_read(tmp_path / "bad.bin")
    (tmp_path / "short.bin").write_bytes(raw[:-10])
    with pytest.raises(L.PSMError):
        L.checkpoint_read(tmp_path / "short.bin")
