"""PhysUnits3D of the C++ facade (the exchange's unit conversions,
src/physunits.h:29-92) with non-identity units, against the physical
meaning of each conversion: with the physical lattice spacing
dx = l_p / N and time step dt = dx u_LB / u_p, a lattice position is x/dx
- 1/2 (cell i centred at (i + 1/2) dx), a velocity u dt/dx, a frequency
w dt, an acceleration a dt^2/dx, a force F rho_p dx^4/dt^2 (kg m/s^2) and a
torque T rho_p dx^5/dt^2; the lattice viscosity is nu_p dt/dx^2, so
tau = 3 nu_p dt/dx^2 + 1/2.  A dropped exponent or factor in the facade
changes one of these.  Host only: the facade header is compiled with g++."""
import json
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "lbdemcoupling-public_amd")

SRC = r'''
#include <cstdio>
#include "facade/plb_ib.h"
using namespace plb;
int main() {
    const double l_p = 0.002, u_p = 0.05, nu_p = 4e-6, rho_p = 950.0, uMax = 0.02;
    const plint N = 12;
    PhysUnits3D<double> U(l_p, u_p, nu_p, 0.02, 0.03, 0.05, N, uMax, rho_p);
    const IncomprFlowParam<double>& P = U.getLbParam();
    const double x = 0.0137, u = 0.0031, w = 2.5, a = 9.81, F = 3e-4, T = -2e-3, t = 0.37;
    std::printf("{\"pos\": %.17g, \"len\": %.17g, \"vel\": %.17g, \"freq\": %.17g, \"accel\": %.17g,"
                " \"force\": %.17g, \"torque\": %.17g, \"time\": %.17g, \"steps\": %lld, \"tau\": %.17g,"
                " \"re\": %.17g, \"nx\": %lld, \"ny\": %lld, \"nz\": %lld,"
                " \"phys_pos\": %.17g, \"phys_vel\": %.17g, \"phys_time\": %.17g,"
                " \"press\": %.17g, \"rho_back\": %.17g, \"dens\": %.17g, \"phys_accel\": %.17g,",
                U.getLbPosition(x), U.getLbLength(x), U.getLbVel(u), U.getLbFreq(w), U.getLbAccel(a),
                U.getPhysForce(F), U.getPhysTorque(T), U.getLbTime(t), (long long)U.getLbSteps(t), P.getTau(),
                U.getRe(), (long long)P.getNx(), (long long)P.getNy(), (long long)P.getNz(),
                U.getPhysPosition(U.getLbPosition(x)), U.getPhysVel(U.getLbVel(u)),
                U.getPhysTime(U.getLbTime(t)), U.getPhysPress(1.003), U.getLbRho(U.getPhysPress(1.003)),
                U.getLbDensity(1200.0), U.getPhysAccel(U.getLbAccel(a)));
    U.setLbOffset(0.25, 0.0, 1.0);
    std::printf(" \"pos_x\": %.17g, \"pos_y\": %.17g, \"pos_z\": %.17g, \"pos_after_offset\": %.17g,"
                " \"phys_pos_z\": %.17g}\n",
                U.getLbPositionX(x), U.getLbPositionY(x), U.getLbPositionZ(x), U.getLbPosition(x),
                U.getPhysPositionZ(U.getLbPositionZ(x)));
    return 0;
}
'''


@pytest.fixture(scope="module")
def units(tmp_path_factory):
    d = tmp_path_factory.mktemp("units")
    src = d / "units.cpp"
    src.write_text(SRC)
    exe = d / "units"
    lib = os.path.join(PKG, "lib")
    cmd = ["g++", "-O1", "-std=c++17", "-I" + PKG, "-I" + os.path.join(ROOT, "include"), str(src), "-o", str(exe),
           "-L" + lib, "-llbdem_psm", "-Wl,-rpath," + lib]
    subprocess.run(cmd, check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    return json.loads(out)


L_P, U_P, NU_P, RHO_P, UMAX, N = 0.002, 0.05, 4e-6, 950.0, 0.02, 12
DX = L_P / N                    # m per cell
DT = DX * UMAX / U_P            # s per step


def test_positions_lengths_and_time(units):
    x, t = 0.0137, 0.37
    assert units["pos"] == pytest.approx(x / DX - 0.5, rel=1e-14)
    assert units["len"] == pytest.approx(x / DX, rel=1e-14)
    assert units["time"] == pytest.approx(t / DT, rel=1e-14)
    assert units["steps"] == int(np.floor(t / DT + 0.5))
    assert units["phys_pos"] == pytest.approx(x, rel=1e-14)
    assert units["phys_time"] == pytest.approx(t, rel=1e-14)
    # setLbOffset moves the per-axis forms only; the exchange keeps 1/2
    assert units["pos_x"] == pytest.approx(x / DX - 0.25, rel=1e-14)
    assert units["pos_y"] == pytest.approx(x / DX, rel=1e-14)
    assert units["pos_z"] == pytest.approx(x / DX - 1.0, rel=1e-14)
    assert units["pos_after_offset"] == pytest.approx(x / DX - 0.5, rel=1e-14)
    assert units["phys_pos_z"] == pytest.approx(x, rel=1e-14)


def test_kinematic_conversions(units):
    assert units["vel"] == pytest.approx(0.0031 * DT / DX, rel=1e-14)
    assert units["phys_vel"] == pytest.approx(0.0031, rel=1e-14)
    assert units["freq"] == pytest.approx(2.5 * DT, rel=1e-14)
    assert units["accel"] == pytest.approx(9.81 * DT * DT / DX, rel=1e-14)
    assert units["phys_accel"] == pytest.approx(9.81, rel=1e-14)


def test_force_and_torque_factors(units):
    assert units["force"] == pytest.approx(3e-4 * RHO_P * DX ** 4 / DT ** 2, rel=1e-13)
    assert units["torque"] == pytest.approx(-2e-3 * RHO_P * DX ** 5 / DT ** 2, rel=1e-13)


def test_pressure_density_and_flow_parameters(units):
    # pressure = (rho - 1) c_s^2 in rho_p (dx/dt)^2
    assert units["press"] == pytest.approx(0.003 / 3.0 * RHO_P * (DX / DT) ** 2, rel=1e-13)
    # getLbRho gives the density deviation rho - 1 of a pressure (physunits.h:62-63), not
    # the inverse of getPhysPress (:76-77), which takes rho: kept as the reference has it
    assert units["rho_back"] == pytest.approx(0.003, rel=1e-12)
    assert units["dens"] == pytest.approx(1200.0 / RHO_P, rel=1e-14)
    assert units["re"] == pytest.approx(L_P * U_P / NU_P, rel=1e-14)
    assert units["tau"] == pytest.approx(3.0 * NU_P * DT / DX ** 2 + 0.5, rel=1e-13)
    assert (units["nx"], units["ny"], units["nz"]) == (121, 181, 301)
