"""Pins of the differentiable pieces of the oracle (CPU only).

* The analytic overlap (orc_overlap_fraction) against Monte Carlo volumes of
  the two-sphere intersection, its closed-form limits, its derivative
  against central differences, and the lens formula's symmetry in the two
  radii.
* The torch FP64 restatement of the frozen-geometry window
  (oracle/psm_torch.py) against the C oracle on the same inputs: records,
  populations and per-step forces/torques within 1e-12, in both
  voxelization modes, with walls, periodic images and Guo forcing.
* torch.autograd of that restatement against central differences, so that
  it can serve as the reference gradient of the HIP adjoint.
"""
import math

import numpy as np
import pytest
import torch

from oracle import oracle as O
from oracle import psm_torch as PT

RC = (3.0 / (4.0 * math.pi)) ** (1.0 / 3.0)


# --------------------------------------------------------------------------- overlap

@pytest.mark.parametrize("r,d", [(5.0, 5.0), (5.0, 4.7), (5.0, 5.5), (0.3, 0.5), (3.0, 2.5), (8.0, 8.3)])
def test_overlap_matches_monte_carlo(r, d):
    rng = np.random.default_rng(int(100 * r + 10 * d))
    n = 2_000_000
    p = rng.normal(size=(n, 3))
    p /= np.linalg.norm(p, axis=1)[:, None]
    p *= RC * rng.random(n)[:, None] ** (1.0 / 3.0)     # uniform in the cell sphere
    inside = np.linalg.norm(p - np.array([d, 0.0, 0.0]), axis=1) < r
    sf, _, br = O.overlap_fraction(d, r)
    assert br == 3
    se = math.sqrt(sf * (1 - sf) / n)
    assert abs(sf - inside.mean()) < 5 * se + 1e-12


def test_overlap_limits():
    # cell fully inside / outside
    assert O.overlap_fraction(5.0 - RC - 1e-9, 5.0)[:3:2] == (1.0, 1)
    assert O.overlap_fraction(5.0 + RC + 1e-9, 5.0)[:3:2] == (0.0, 0)
    # a sphere inside the cell sphere covers its own volume
    sf, der, br = O.overlap_fraction(0.1, 0.2)
    assert br == 4 and der == 0.0
    assert sf == pytest.approx((4 / 3) * math.pi * 0.2 ** 3, rel=1e-14)
    # continuity at both ends of the lens branch
    for r in (0.3, 5.0):
        lo, hi = abs(r - RC), r + RC
        assert O.overlap_fraction(lo * (1 + 1e-10), r)[0] == pytest.approx(O.overlap_fraction(lo, r)[0], abs=1e-8)
        assert O.overlap_fraction(hi * (1 - 1e-10), r)[0] == pytest.approx(0.0, abs=1e-8)
    # the lens volume is symmetric in the two radii: V(d; r, RC) = V(d; RC, r)
    # with the roles swapped, checked through the unit-volume normalisation
    r, d = 0.7, 0.9
    t = (r + RC) - d
    v_swapped = math.pi * t * t * (d * d + 2 * d * (RC + r) - 3 * (RC - r) ** 2) / (12 * d)
    assert O.overlap_fraction(d, r)[0] == pytest.approx(v_swapped / ((4 / 3) * math.pi * RC ** 3), rel=1e-13)


@pytest.mark.parametrize("r", [0.3, 2.0, 5.0, 9.5])
def test_overlap_derivative_central_difference(r):
    for d in np.linspace(abs(r - RC) + 0.01, r + RC - 0.01, 9):
        _, der, _ = O.overlap_fraction(d, r)
        e = 1e-6
        fd = (O.overlap_fraction(d + e, r)[0] - O.overlap_fraction(d - e, r)[0]) / (2 * e)
        assert der == pytest.approx(fd, rel=1e-6, abs=1e-9)


def test_overlap_torch_matches_c():
    r = torch.tensor([5.0] * 7 + [0.3] * 3, dtype=torch.float64)
    d = torch.tensor([0.0, 3.0, 4.5, 5.0, 5.4, 5.7, 6.0, 0.1, 0.5, 1.0], dtype=torch.float64)
    got = PT.overlap(d, r).numpy()
    want = [O.overlap_fraction(float(a), float(b))[0] for a, b in zip(d, r)]
    np.testing.assert_allclose(got, want, rtol=1e-14, atol=1e-16)


# --------------------------------------------------------------------------- window

CASE = dict(n=(14, 12, 16), periodic=(0, 1, 1), tau=0.8, body=(0.0, 0.0, 2e-5))


def spheres():
    x = np.array([[6.3, 5.2, 3.1], [7.6, 4.1, 10.4], [4.2, 8.7, 14.6]])   # third straddles z
    v = np.array([[0.01, -0.004, 0.003], [-0.006, 0.002, 0.008], [0.003, 0.005, -0.007]])
    w = np.array([[0.001, 0.002, -0.003], [0.0, -0.002, 0.001], [0.002, 0.0, 0.001]])
    r = np.array([3.1, 2.7, 2.4])
    ids = np.array([1, 2, 3], dtype=np.int64)
    return x, v, w, r, ids


def oracle_run(mode, nsteps, seed=3):
    n = CASE["n"]
    o = O.OracleLattice(*n, periodic=CASE["periodic"], tau=CASE["tau"], slices_per_dim=5,
                        body_accel=CASE["body"], voxel_mode=1 if mode == "overlap" else 0)
    rng = np.random.default_rng(seed)
    nc = int(np.prod(n))
    o.init_equilibrium(1 + rng.normal(0, 0.01, nc), rng.normal(0, 0.02, (3, nc)))
    f0 = o.get_populations()
    x, v, w, r, ids = spheres()
    o.set_spheres(x, v, w, r, ids)
    rec = o.cell_records()
    fts = []
    for _ in range(nsteps):
        o.collide_stream(1, 1)
        F, T = o.sum_force_torque(x)
        fts.append(np.concatenate([F, T], 1))
    return f0, rec, np.array(fts), o.get_populations()


def frozen(rec):
    live = rec["sf"] > 0
    owner = np.where(live, rec["id"] - 1, 0)     # ids 1..n map to index id-1
    return PT.FrozenWindow(CASE["n"], CASE["periodic"], live, owner, acc=CASE["body"], sf0=rec["sf"])


def torch_params():
    x, v, w, r, _ = spheres()
    t = lambda a: torch.tensor(a, dtype=torch.float64, requires_grad=True)
    return t(x), t(v), t(w), torch.tensor(r, dtype=torch.float64), t(CASE["tau"])


@pytest.mark.parametrize("mode", ["subcell", "overlap"])
def test_torch_window_matches_oracle(mode):
    nsteps = 6
    f0, rec, ft_o, f_o = oracle_run(mode, nsteps)
    fw = frozen(rec)
    x, v, w, r, tau = torch_params()
    dsf = torch.zeros(fw.nc, dtype=torch.float64)
    # the record values the torch restatement derives from the parameters
    d = fw.arms(x.detach())
    us = (v.detach()[fw.lo] + torch.cross(w.detach()[fw.lo], d, dim=1)).numpy()
    np.testing.assert_allclose(us, rec["us"][:, fw.lc.numpy()].T, rtol=0, atol=1e-15)
    if mode == "overlap":
        sf = PT.overlap(torch.sqrt((d * d).sum(1)), r[fw.lo]).numpy()
        np.testing.assert_allclose(sf, rec["sf"][fw.lc.numpy()], rtol=1e-13, atol=1e-15)
    ft, f = fw.run(torch.tensor(f0.T.copy()), x, v, w, r, tau, dsf, nsteps, mode)
    ft = ft.detach().numpy()
    scale = np.abs(ft_o).max()
    assert scale > 1e-6
    assert np.abs(ft - ft_o).max() / scale < 1e-12
    assert np.abs(f.detach().numpy().T - f_o).max() < 1e-14


@pytest.mark.parametrize("mode", ["subcell", "overlap"])
def test_torch_autograd_matches_central_differences(mode):
    nsteps = 4
    f0, rec, _, _ = oracle_run(mode, 0)
    fw = frozen(rec)
    x, v, w, r, tau = torch_params()
    dsf = torch.zeros(fw.nc, dtype=torch.float64, requires_grad=True)
    f0t = torch.tensor(f0.T.copy())
    rng = np.random.default_rng(7)
    g = torch.tensor(rng.normal(size=(nsteps, 3, 6)))

    def J(x_, v_, w_, tau_, dsf_):
        ft, _ = fw.run(f0t, x_, v_, w_, r, tau_, dsf_, nsteps, mode)
        return (ft * g).sum()

    J(x, v, w, tau, dsf).backward()
    # a partial cell of particle 1 for the solid-fraction direction
    sfl = rec["sf"][fw.lc.numpy()]
    k = int(np.argmin(np.abs(sfl - 0.5)))
    cell = int(fw.lc[k])
    checks = [("x", x, (0, 1)), ("x", x, (2, 2)), ("v", v, (1, 0)), ("w", w, (0, 2)),
              ("tau", tau, ()), ("dsf", dsf, (cell,))]
    for name, p, ix in checks:
        eps = 1e-6
        base = [x.detach(), v.detach(), w.detach(), tau.detach(), dsf.detach()]
        which = ["x", "v", "w", "tau", "dsf"].index(name)
        plus = [b.clone() for b in base]
        minus = [b.clone() for b in base]
        plus[which][ix] += eps
        minus[which][ix] -= eps
        fd = (J(*plus) - J(*minus)).item() / (2 * eps)
        ad = p.grad[ix].item()
        assert ad == pytest.approx(fd, rel=2e-6, abs=1e-12), (name, ix, ad, fd)
