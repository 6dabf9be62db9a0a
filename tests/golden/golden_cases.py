"""Small seeded cases of the coupled PSM step, and how to replay them.

The reference (LBDEMcoupling on Palabos and LIGGGHTS) ships no tests, golden
vectors or fixtures, and it cannot be built here.  The fixtures in this
folder are therefore the FP64 oracle's outputs on these cases, written once
by make_golden.py.  They pin the oracle and the HIP path against drift:
tests/test_golden.py replays them through the oracle (bit for bit) and
tests/test_gpu_golden.py through the C ABI (bit for bit for populations,
records and sub-cell counts; forces and torques within 1e-12).

Each case is the reference driver's loop (examples/sedimentingSphere/
sedimentingSphere.cpp:212-225): setSpheresOnLattice, collideAndStream,
getForcesFromLattice, with prescribed kinematics x += v between steps.
Positions of every step are stored in the fixture, so a replay needs no
kinematics of its own.
"""
from __future__ import annotations

import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
Q = 19

# name -> parameters.  init: "eq" (random rho/u field) or "poiseuille"
# (PoiseuilleProfileAndPressureGradient with the Zhang-Kwok periodic pressure).
CASES = {
    # fully periodic, N = 5, overlapping spheres (the four-case ownership
    # rule with shared cells) and spheres across the x and y faces (images)
    "tiny_periodic": dict(n=(9, 8, 10), periodic=(1, 1, 1), tau=0.7, N=5, body=(0.0, 0.0, 0.0),
                          init="eq", nsph=4, rlo=2.2, rhi=3.2, sigma_v=0.01, sigma_w=1e-3,
                          steps=4, init_vel=False, faces=(0, 1), seed=101),
    # walls on x and y (half-way bounce-back), Guo body force, N = 8, the
    # equilibrium initialisation of solid cells on the first step (initVel)
    "tiny_walls": dict(n=(10, 9, 8), periodic=(0, 0, 1), tau=0.9, N=8, body=(0.0, 1e-6, 2e-6),
                       init="eq", nsph=3, rlo=2.0, rhi=2.6, sigma_v=0.008, sigma_w=1e-3,
                       steps=5, init_vel=True, faces=(2,), seed=202),
    # pressure-driven channel along x (periodic x, walls y and z)
    "tiny_channel": dict(n=(12, 8, 9), periodic=(1, 0, 0), tau=0.6, N=5, body=(0.0, 0.0, 0.0),
                         init="poiseuille", rho_in=1.001, rho_out=0.999, u_max=0.02,
                         nsph=2, rlo=2.0, rhi=2.5, sigma_v=0.005, sigma_w=0.0,
                         steps=6, init_vel=False, faces=(0,), seed=303),
}

# sub-cell count queries: N -> (count, seed)
COUNT_CASES = {5: (2500, 11), 8: (2500, 12)}


def fixture_path(name: str) -> str:
    return os.path.join(HERE, name + ".npz")


def _advance(x, v, n, periodic):
    x = x + v
    for d in range(3):
        if periodic[d]:
            L = float(n[d])
            col = x[:, d]
            col[col >= L - 0.5] -= L
            col[col < -0.5] += L
    return x


def make_inputs(name: str) -> dict:
    """The case's inputs from its seed (numpy only)."""
    c = CASES[name]
    n, per = c["n"], c["periodic"]
    rng = np.random.default_rng(c["seed"])
    nc = int(np.prod(n))
    k = c["nsph"]
    r = rng.uniform(c["rlo"], c["rhi"], k)
    x = np.empty((k, 3))
    for d in range(3):
        if per[d]:
            x[:, d] = rng.uniform(-0.5, n[d] - 0.5, k)
        else:
            # clear of the walls by a cell beyond the radius
            x[:, d] = rng.uniform(r + 1.0, n[d] - 2.0 - r)
    # push one sphere across each listed periodic face
    for s, d in enumerate(c["faces"]):
        if per[d] and s < k:
            x[s, d] = -0.3 if s % 2 == 0 else n[d] - 0.6
    v = rng.normal(0.0, c["sigma_v"], (k, 3))
    w = rng.normal(0.0, c["sigma_w"], (k, 3)) if c["sigma_w"] > 0 else np.zeros((k, 3))
    for d in range(3):
        if not per[d]:
            v[:, d] = 0.0   # stay clear of the walls
    ids = np.arange(1, k + 1, dtype=np.int64) * 3   # tags need not be 1..k
    xs = [x]
    for _ in range(1, c["steps"]):
        xs.append(_advance(xs[-1], v, n, per))
    inp = dict(x_steps=np.array(xs), v=v, w=w, r=r, ids=ids)
    if c["init"] == "eq":
        inp["rho"] = 1.0 + rng.normal(0.0, 0.01, nc)
        inp["u"] = rng.normal(0.0, 0.01, (3, nc))
    return inp


def replay(lat, name: str, inp: dict) -> dict:
    """Run the case on a lattice with the common method names
    (lbdem_psm.Lattice or oracle.OracleLattice) and return its outputs."""
    c = CASES[name]
    n = c["n"]
    if c["init"] == "eq":
        lat.init_equilibrium(inp["rho"], inp["u"])
    else:
        d = [i for i in range(3) if c["periodic"][i]][0]
        n3 = list(n)
        n3[d] -= 1
        lat.init_poiseuille(d, c["rho_in"], c["rho_out"], c["u_max"], n3=n3)
        lat.set_periodic_pressure(d, c["rho_in"], c["rho_out"])
    Fs, Ts = [], []
    for s in range(c["steps"]):
        x = inp["x_steps"][s]
        lat.set_spheres(x, inp["v"], inp["w"], inp["r"], inp["ids"], init_vel=(c["init_vel"] and s == 0))
        lat.collide_stream(1)
        F, T = lat.sum_force_torque(x)
        Fs.append(F)
        Ts.append(T)
    rec = lat.cell_records()
    return dict(f=lat.get_populations(), sf=rec["sf"], id=rec["id"], us=rec["us"], F_cell=rec["F"],
                force=np.array(Fs), torque=np.array(Ts))


def make_count_queries(N: int):
    """Offsets and radii of the sub-cell count queries: centres near the
    sphere surface (inside the shortcut band), a fifth on an exact 1/8 grid."""
    m, seed = COUNT_CASES[N]
    rng = np.random.default_rng(seed)
    r = rng.uniform(2.5, 10.0, m)
    d = rng.normal(0.0, 1.0, (m, 3))
    d *= ((r + rng.uniform(-1.0, 1.0, m)) / np.linalg.norm(d, axis=1))[:, None]
    q = m // 5
    d[:q] = np.round(d[:q] * 8) / 8
    return d, r


def load_fixture(name: str) -> dict:
    with np.load(fixture_path(name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


INPUT_KEYS = ("x_steps", "v", "w", "r", "ids", "rho", "u")
