"""The HIP path (through the C ABI) against the oracle's committed fixtures
(tests/golden/): populations, records, per-cell forces and sub-cell counts
bit for bit, per-particle force and torque within 1e-12 (only the order of
the sums differs)."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))

import golden_cases as G  # noqa: E402
import lbdem_psm as L  # noqa: E402

pytestmark = pytest.mark.gpu


def rel_err(a, b):
    den = np.abs(b).max()
    return float(np.abs(a - b).max() / den) if den > 0 else float(np.abs(a - b).max())


@pytest.mark.parametrize("name", list(G.CASES))
def test_hip_path_reproduces_fixture(name):
    z = G.load_fixture(name)
    c = G.CASES[name]
    print("case", name, c["n"], flush=True)
    g = L.Lattice(*c["n"], periodic=c["periodic"], tau=c["tau"], precision="fp64",
                  slices_per_dim=c["N"], body_accel=c["body"])
    inp = {k: z[k] for k in G.INPUT_KEYS if k in z}
    out = G.replay(g, name, inp)
    g.close()
    for k in ("f", "sf", "id", "us", "F_cell"):
        assert np.array_equal(out[k], z[k]), (name, k)
    for s in range(c["steps"]):
        F, T, Fo, To = out["force"][s], out["torque"][s], z["force"][s], z["torque"][s]
        assert rel_err(F, Fo) < 1e-12, (name, s)
        den = max(np.abs(To).max(), 10.0 * np.abs(Fo).max())
        assert np.abs(T - To).max() / den < 1e-12, (name, s)


@pytest.mark.parametrize("N", list(G.COUNT_CASES))
def test_hip_subcell_counts_reproduce_fixture(N):
    z = G.load_fixture(f"subcell_counts_N{N}")
    with L.Lattice(4, 4, 4, slices_per_dim=N) as g:
        sf, cnt, br = g.subcell_counts(z["d"], z["r"])
    assert np.array_equal(cnt, z["count"])
    assert np.array_equal(br, z["branch"])
    assert np.array_equal(sf, z["sf"])
