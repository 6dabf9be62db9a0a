"""C ABI checks that need no GPU: the library loads, exports every function
declared in include/*.h, and rejects bad descriptors / particle arrays on the
host with the documented error codes (never an abort)."""
import ctypes
import os
import re

import numpy as np
import pytest

import lbdem_psm as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    inc = os.path.join(ROOT, "include")
    for fn in sorted(os.listdir(inc)):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(inc, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"\b(lbdem_psm_\w+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = L.load_library()
    names = declared_functions()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n
    assert b"gfx950" in lib.lbdem_psm_version()


def test_python_binding_covers_abi():
    assert set(declared_functions()) <= set(L._SIGS)


def test_desc_layout_matches_header():
    # nx,ny,nz,periodic[3] (24) + tau (8) + precision, slices (8) + accel (24) + device,
    # voxel_mode (8) + max_solid_cells (8)
    assert ctypes.sizeof(L.Desc) == 80


def test_desc_validation():
    d = L.make_desc(8, 8, 8)
    assert L.check_desc(d) == L.OK
    for field, val, code in [("tau", 0.5, L.ERR_NUMERIC), ("tau", float("nan"), L.ERR_NUMERIC),
                             ("nx", 0, L.ERR_INVALID_ARG), ("slices_per_dim", 0, L.ERR_INVALID_ARG),
                             ("slices_per_dim", 17, L.ERR_INVALID_ARG), ("precision", 7, L.ERR_INVALID_ARG),
                             ("max_solid_cells", 10 ** 6, L.ERR_INVALID_ARG), ("voxel_mode", 2, L.ERR_INVALID_ARG)]:
        d = L.make_desc(8, 8, 8)
        setattr(d, field, val)
        assert L.check_desc(d) == code, field
    d = L.make_desc(2048, 2048, 1024)
    assert L.check_desc(d) == L.ERR_INVALID_ARG  # more than 2^31 - 1 cells


def test_sphere_validation():
    d = L.make_desc(16, 16, 16)
    x = np.array([[4.0, 4.0, 4.0]])
    z = np.zeros((1, 3))
    assert L.check_spheres(d, x, z, z, [2.0], [1]) == L.OK
    assert L.check_spheres(d, x, z, z, [0.0], [1]) == L.ERR_INVALID_ARG
    assert L.check_spheres(d, x, z, z, [2.0], [0]) == L.ERR_INVALID_ARG
    assert L.check_spheres(d, [[40.0, 4, 4]], z, z, [2.0], [1]) == L.ERR_INVALID_ARG
    assert L.check_spheres(d, [[np.nan, 4, 4]], z, z, [2.0], [1]) == L.ERR_NUMERIC
    assert L.check_spheres(d, x, [[0.31, 0, 0]], z, [2.0], [1]) == L.ERR_NUMERIC


def test_create_without_device_reports_error():
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    with pytest.raises(L.PSMError) as e:
        L.Lattice(8, 8, 8)
    assert e.value.code == L.ERR_NO_DEVICE


def test_workload_generator_is_seeded_and_non_overlapping():
    box = (64, 64, 64)
    a = L.generate_spheres(box, 100, L.RADIUS_UNIFORM, a=3, b=6, sigma_v=0.01, sigma_w=1e-3, seed=9)
    b = L.generate_spheres(box, 100, L.RADIUS_UNIFORM, a=3, b=6, sigma_v=0.01, sigma_w=1e-3, seed=9)
    c = L.generate_spheres(box, 100, L.RADIUS_UNIFORM, a=3, b=6, sigma_v=0.01, sigma_w=1e-3, seed=10)
    for p, q in zip(a, b):
        assert np.array_equal(p, q)
    assert not np.array_equal(a[0], c[0])
    x, v, w, r, ids = a
    assert np.all((x >= -0.5) & (x < 63.5))
    assert np.all((r >= 3) & (r <= 6))
    assert np.array_equal(ids, np.arange(1, 101))
    d = x[:, None, :] - x[None, :, :]
    d -= 64 * np.round(d / 64)
    dist = np.sqrt((d ** 2).sum(-1)) + np.eye(100) * 1e9
    assert np.all(dist >= r[:, None] + r[None, :] - 1e-12)
    assert abs(v.std() - 0.01) < 0.002 and abs(w.std() - 1e-3) < 2e-4


def test_lognormal_radii_clipped():
    x, v, w, r, ids = L.generate_spheres((96, 96, 96), 200, L.RADIUS_LOGNORMAL, a=5.0, b=0.25, lo=3.0,
                                         hi=10.0, seed=3)
    assert r.min() >= 3.0 and r.max() <= 10.0
    assert abs(np.median(r) - 5.0) < 0.3


def test_advance_positions_wraps():
    x = np.array([[63.45, -0.45, 10.0]])
    v = np.array([[0.1, -0.1, 0.0]])
    y = L.advance_positions(x, v, (64, 64, 64), (1, 1, 0))
    assert y[0, 0] == (63.45 + 0.1) - 64.0
    assert y[0, 1] == (-0.45 - 0.1) + 64.0
