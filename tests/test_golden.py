"""The oracle against its committed fixtures (tests/golden/).

The reference ships no golden vectors; these fixtures are the oracle's own
outputs on small seeded cases (tests/golden/make_golden.py), so this pins
the oracle against drift: any change to its arithmetic, its ownership rule,
its streaming or its sums shows up here bit for bit.  What pins the oracle
to the reference is in test_oracle.py.
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))

import golden_cases as G  # noqa: E402
from oracle import oracle as O  # noqa: E402


def test_fixtures_hold_the_generated_inputs():
    """The stored inputs are the ones the case's seed generates."""
    for name in G.CASES:
        z = G.load_fixture(name)
        inp = G.make_inputs(name)
        for k, v in inp.items():
            assert np.array_equal(z[k], v), (name, k)


@pytest.mark.parametrize("name", list(G.CASES))
def test_oracle_reproduces_fixture(name):
    z = G.load_fixture(name)
    c = G.CASES[name]
    o = O.OracleLattice(*c["n"], periodic=c["periodic"], tau=c["tau"], slices_per_dim=c["N"],
                        body_accel=c["body"])
    o.set_threads(1)
    inp = {k: z[k] for k in G.INPUT_KEYS if k in z}
    out = G.replay(o, name, inp)
    o.close()
    for k in ("f", "sf", "id", "us", "F_cell", "force", "torque"):
        assert np.array_equal(out[k], z[k]), (name, k)
    # the case exercises what it is meant to: solid cells and a non-zero force
    assert np.count_nonzero(z["sf"] > 1e-3) > 50
    assert np.abs(z["force"]).max() > 0


@pytest.mark.parametrize("N", list(G.COUNT_CASES))
def test_oracle_subcell_counts_reproduce_fixture(N):
    z = G.load_fixture(f"subcell_counts_N{N}")
    for k in range(len(z["r"])):
        sf, cnt, br = O.calc_solid_fraction(*z["d"][k], z["r"][k], N)
        assert (sf, cnt, br) == (z["sf"][k], z["count"][k], z["branch"][k]), k
    # the fixture holds partial cells, and the sf = count / N^3 identity
    assert np.count_nonzero((z["count"] > 0) & (z["count"] < N ** 3)) > 1000
    assert np.array_equal(z["sf"], z["count"] * (1.0 / N ** 3))
