"""The C++ facade (facade/plb_ib.h) drives the library like the reference's
sedimentingSphere driver; run the built example for a few hundred steps."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "lbdemcoupling-public_amd", "examples", "sedimenting_sphere")


def test_facade_example_builds():
    assert os.path.exists(EXE), "make -C lbdemcoupling-public_amd builds the example"


@pytest.mark.gpu
def test_facade_sedimenting_sphere_runs():
    out = subprocess.run([EXE, "400", "100"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    rows = re.findall(r"step (\d+) z ([-\d.e+]+) vz ([-\d.e+]+) fz ([-\d.e+]+)", out.stdout)
    assert len(rows) == 4
    z = [float(r[1]) for r in rows]
    vz = [float(r[2]) for r in rows]
    fz = [float(r[3]) for r in rows]
    # the sphere accelerates downwards from rest, the fluid pushes back up
    assert z[-1] < 100.0 and all(v < 0 for v in vz) and fz[-1] > 0
    assert vz[-1] < vz[0]


EXE2 = os.path.join(ROOT, "lbdemcoupling-public_amd", "examples", "rect_channel")


def test_facade_rect_channel_builds():
    assert os.path.exists(EXE2), "make -C lbdemcoupling-public_amd builds the example"


@pytest.mark.gpu
def test_facade_rect_channel_runs(tmp_path):
    """Pressure-driven channel through the facade: the flow goes along +x,
    the spheres feel a drag along +x, the VTK file parses and a checkpoint
    continues bit for bit."""
    prefix = str(tmp_path / "rc")
    out = subprocess.run([EXE2, "300", prefix], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    rows = re.findall(r"step (\d+) flux ([-\d.e+]+) fx ([-\d.e+]+)", out.stdout)
    assert len(rows) == 3
    flux = [float(r[1]) for r in rows]
    fx = [float(r[2]) for r in rows]
    assert all(f > 0 for f in flux) and fx[-1] > 0
    assert "checkpoint ok" in out.stdout
    from test_io import read_vti
    n, _, fields = read_vti(prefix + ".vti")
    assert n == (48, 24, 24)
    assert set(fields) == {"velocity", "density", "pressure", "solidfraction"}
    assert fields["solidfraction"].max() == 1.0
