"""Write the oracle's fixtures for the cases of golden_cases.py.

    python tests/golden/make_golden.py        (from the repository root)

Needs the oracle built (make -C oracle).  Each case's fixture holds its
inputs (positions of every step, velocities, spin, radii, tags, the initial
field) and the oracle's outputs after the last step (populations, records,
per-cell forces) with the per-particle force and torque of every step.  The
oracle runs sequentially (one x-slab), so its sums are in the reference's
lexicographic order.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

from oracle import oracle as O  # noqa: E402
import golden_cases as G  # noqa: E402


def oracle_lattice(name):
    c = G.CASES[name]
    o = O.OracleLattice(*c["n"], periodic=c["periodic"], tau=c["tau"], slices_per_dim=c["N"],
                        body_accel=c["body"])
    o.set_threads(1)
    return o


def main():
    for name in G.CASES:
        inp = G.make_inputs(name)
        o = oracle_lattice(name)
        out = G.replay(o, name, inp)
        o.close()
        np.savez_compressed(G.fixture_path(name), **inp, **out)
        live = int(np.count_nonzero(out["sf"] > 1e-3))
        print(f"{name}: {G.CASES[name]['steps']} steps, {len(inp['r'])} spheres, {live} solid cells, "
              f"max |F| {np.abs(out['force']).max():.3e}")
    for N in G.COUNT_CASES:
        d, r = G.make_count_queries(N)
        res = [O.calc_solid_fraction(*d[k], r[k], N) for k in range(len(r))]
        sf = np.array([t[0] for t in res])
        cnt = np.array([t[1] for t in res], dtype=np.int32)
        br = np.array([t[2] for t in res], dtype=np.int32)
        np.savez_compressed(G.fixture_path(f"subcell_counts_N{N}"), d=d, r=r, sf=sf, count=cnt, branch=br)
        print(f"subcell_counts_N{N}: {len(r)} queries, {int(np.count_nonzero((cnt > 0) & (cnt < N ** 3)))} partial")


if __name__ == "__main__":
    main()
