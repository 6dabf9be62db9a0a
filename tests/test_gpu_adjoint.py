"""The differentiable window on the GPU: forward parity of the analytic-overlap
voxelization, and the hand-written HIP adjoint against

* the FP64 reference gradient: torch.autograd of the torch restatement
  (oracle/psm_torch.py, itself pinned to the C oracle), 1e-9 relative;
* central finite differences (eps 1e-6, FP64) of the GPU forward itself,
  through the torch.autograd.Function, on 20 random parameters, 1e-5;
* the transpose identity <J u, w> = <u, J^T w> of one collide-stream step;
* checkpoint-interval invariance (bit-identical gradients for every K);
* the FP32-deviation adjoint against the FP64 gradient.
"""
import numpy as np
import pytest
import torch

import lbdem_psm as L
from lbdem_psm.autograd import psm_window
from oracle import oracle as O
from oracle import psm_torch as PT

pytestmark = pytest.mark.gpu

CASE = dict(n=(14, 12, 16), periodic=(0, 1, 1), tau=0.8, body=(0.0, 0.0, 2e-5))


def spheres():
    x = np.array([[6.3, 5.2, 3.1], [7.6, 4.1, 10.4], [4.2, 8.7, 14.6]])
    v = np.array([[0.01, -0.004, 0.003], [-0.006, 0.002, 0.008], [0.003, 0.005, -0.007]])
    w = np.array([[0.001, 0.002, -0.003], [0.0, -0.002, 0.001], [0.002, 0.0, 0.001]])
    r = np.array([3.1, 2.7, 2.4])
    ids = np.array([1, 2, 3], dtype=np.int64)
    return x, v, w, r, ids


def initial_populations(n, seed=3):
    o = O.OracleLattice(*n, periodic=(1, 1, 1))
    rng = np.random.default_rng(seed)
    nc = int(np.prod(n))
    o.init_equilibrium(1 + rng.normal(0, 0.01, nc), rng.normal(0, 0.02, (3, nc)))
    return o.get_populations()


def gpu_lattice(mode, precision="fp64"):
    return L.Lattice(*CASE["n"], periodic=CASE["periodic"], tau=CASE["tau"], precision=precision,
                     body_accel=CASE["body"], voxel_mode=mode)


def rel(a, b):
    den = np.abs(b).max()
    return float(np.abs(a - b).max() / den) if den > 0 else float(np.abs(a - b).max())


def test_overlap_voxelization_matches_oracle():
    n = CASE["n"]
    x, v, w, r, ids = spheres()
    g = gpu_lattice("overlap")
    o = O.OracleLattice(*n, periodic=CASE["periodic"], tau=CASE["tau"], body_accel=CASE["body"],
                        voxel_mode=1)
    f0 = initial_populations(n)
    g.set_populations(f0)
    o.set_populations(f0)
    g.set_spheres(x, v, w, r, ids)
    o.set_spheres(x, v, w, r, ids)
    rg, ro = g.cell_records(), o.cell_records()
    assert np.array_equal(rg["id"], ro["id"])
    assert np.array_equal(rg["us"], ro["us"])
    assert np.abs(rg["sf"] - ro["sf"]).max() < 1e-14
    assert ((ro["sf"] > 0.001) & (ro["sf"] < 0.999)).sum() > 100
    for _ in range(5):
        g.collide_stream(1)
        o.collide_stream(1, 1)
        Fg, Tg = g.sum_force_torque(x)
        Fo, To = o.sum_force_torque(x)
        assert rel(Fg, Fo) < 1e-12
        assert rel(Tg, To) < 1e-12
    assert rel(g.get_populations(), o.get_populations()) < 1e-13
    # GPU analytic overlap of single queries equals the oracle's
    rng = np.random.default_rng(5)
    rr = rng.uniform(0.3, 9.0, 4000)
    d = rng.normal(size=(4000, 3))
    d *= ((rr + rng.uniform(-0.7, 0.7, 4000)) / np.linalg.norm(d, axis=1))[:, None]
    sf, _, br = g.subcell_counts(d, rr)
    want = np.array([O.overlap_fraction(float(np.sqrt((a[0] * a[0] + a[1] * a[1]) + a[2] * a[2])), float(b))[0]
                     for a, b in zip(d, rr)])
    assert np.abs(sf - want).max() < 1e-14
    assert set(np.unique(br)) <= {0, 1, 3, 4}


def reference_gradient(mode, f0, rec, nsteps, G):
    live = rec["sf"] > 0
    owner = np.where(live, rec["id"] - 1, 0)
    fw = PT.FrozenWindow(CASE["n"], CASE["periodic"], live, owner, acc=CASE["body"], sf0=rec["sf"])
    x, v, w, r, _ = spheres()
    t = lambda a: torch.tensor(a, dtype=torch.float64, requires_grad=True)
    x, v, w, tau = t(x), t(v), t(w), t(CASE["tau"])
    dsf = torch.zeros(fw.nc, dtype=torch.float64, requires_grad=True)
    ft, _ = fw.run(torch.tensor(f0.T.copy()), x, v, w, torch.tensor(r), tau, dsf, nsteps, mode)
    (ft * torch.tensor(G)).sum().backward()
    return ft.detach().numpy(), {"x": x.grad.numpy(), "v": v.grad.numpy(), "w": w.grad.numpy(),
                                 "tau": tau.grad.item(), "sf": dsf.grad.numpy()}


@pytest.mark.parametrize("mode", ["subcell", "overlap"])
def test_adjoint_matches_reference_gradient(mode):
    nsteps, K = 9, 4
    f0 = initial_populations(CASE["n"])
    x, v, w, r, ids = spheres()
    g = gpu_lattice(mode)
    g.set_populations(f0)
    g.set_spheres(x, v, w, r, ids)
    rec = g.cell_records()
    ft = g.diff_forward(nsteps, K)
    G = np.random.default_rng(11).normal(size=ft.shape)
    grads = g.diff_backward(G)
    ft_ref, ref = reference_gradient(mode, f0, rec, nsteps, G)
    assert rel(ft, ft_ref) < 1e-12
    for k in ("x", "v", "w", "sf"):
        assert rel(grads[k], ref[k]) < 1e-9, k
    assert grads["tau"] == pytest.approx(ref["tau"], rel=1e-9)
    # dJ/dB and dJ/dsf are related by dB/dsf on partial cells
    ooo = CASE["tau"] - 0.5
    sf = rec["sf"]
    part = (sf > 0.001) & (sf <= 0.999)
    dBds = ooo * (1 + ooo) / ((1 - sf[part]) + ooo) ** 2
    np.testing.assert_allclose(grads["sf"][part], grads["B"][part] * dBds, rtol=1e-12, atol=1e-300)
    assert np.all(grads["B"][~part] == 0.0)


def test_adjoint_checkpoint_interval_invariance():
    nsteps = 7
    f0 = initial_populations(CASE["n"])
    x, v, w, r, ids = spheres()
    g = gpu_lattice("overlap")
    G = np.random.default_rng(2).normal(size=(nsteps, 3, 6))
    outs = []
    # K > 0: a checkpoint every K steps; K = -M: a budget of M states with
    # segments of M, M-1, ... steps (4 + 3, 5 + 2, 7)
    for K in (1, 3, 7, 10, -4, -5, -7):
        g.set_populations(f0)
        g.clear_records()
        g.set_spheres(x, v, w, r, ids)
        g.diff_forward(nsteps, K)
        outs.append(g.diff_backward(G, want_f0=True))
        if K < 0:
            with pytest.raises(L.PSMError):   # the budgeted window is consumed
                g.diff_backward(G)
    for o in outs[1:]:
        for k in ("x", "v", "w", "sf", "B", "f0"):
            assert np.array_equal(o[k], outs[0][k]), k
        assert o["tau"] == outs[0]["tau"]
    g.diff_release()


@pytest.mark.parametrize("K", [3, -4])
def test_window_forward_leaves_the_lattice_as_plain_steps(K):
    """The window's forward keeps each segment's checkpoint by taking over
    the buffer its first step read: afterwards the lattice holds exactly the
    state, forces and records of the same steps run without a window, and it
    steps on identically; the checkpoints still reproduce the gradients."""
    nsteps = 7
    f0 = initial_populations(CASE["n"])
    x, v, w, r, ids = spheres()
    a, b = gpu_lattice("overlap"), gpu_lattice("overlap")
    try:
        for lat in (a, b):
            lat.set_populations(f0)
            lat.set_spheres(x, v, w, r, ids)
        ft = a.diff_forward(nsteps, K)
        b.step(nsteps, frozen=True)
        assert np.array_equal(a.get_populations(), b.get_populations())
        Fb, Tb = b.particle_forces()
        assert np.array_equal(ft[-1][:, :3], Fb) and np.array_equal(ft[-1][:, 3:], Tb)
        G = np.random.default_rng(5).normal(size=(nsteps, 3, 6))
        g1 = a.diff_backward(G, want_f0=True)
        for lat in (a, b):
            lat.step(3, frozen=True)
        assert np.array_equal(a.get_populations(), b.get_populations())
        # a second window from the stepped state reuses the state pool
        a.set_populations(f0)
        a.diff_forward(nsteps, K)
        g2 = a.diff_backward(G, want_f0=True)
        for k in ("x", "v", "w", "sf", "B", "f0"):
            assert np.array_equal(g1[k], g2[k]), k
    finally:
        a.close()
        b.close()


def test_adjoint_state_budget_too_small():
    x, v, w, r, ids = spheres()
    g = gpu_lattice("overlap")
    g.set_populations(initial_populations(CASE["n"]))
    g.set_spheres(x, v, w, r, ids)
    with pytest.raises(L.PSMError):
        g.diff_forward(7, -3)      # 3 + 2 + 1 < 7 steps
    g.diff_forward(6, -3)          # 3 + 2 + 1 = 6
    g.diff_release()


def test_adjoint_transpose_identity_one_step():
    """<(S(f0 + e u) - S(f0 - e u)) / 2e, w> = <u, S^T w> for one step S."""
    n = CASE["n"]
    f0 = initial_populations(n)
    x, v, w, r, ids = spheres()
    g = gpu_lattice("overlap")
    g.set_populations(f0)
    g.set_spheres(x, v, w, r, ids)
    rng = np.random.default_rng(4)
    u = rng.normal(size=f0.shape)
    wt = rng.normal(size=f0.shape)
    g.diff_forward(1, 1)
    lam0 = g.diff_backward(g_fT=wt, want_f0=True)["f0"]
    eps = 1e-6
    g.set_populations(f0 + eps * u)
    g.diff_forward(1, 1)
    fp = g.get_populations()
    g.set_populations(f0 - eps * u)
    g.diff_forward(1, 1)
    fm = g.get_populations()
    lhs = float(np.sum((fp - fm) / (2 * eps) * wt))
    rhs = float(np.sum(u * lam0))
    assert lhs == pytest.approx(rhs, rel=1e-7)


def pp_lattice(precision="fp64"):
    """CASE with a Zhang-Kwok pressure difference along the periodic z axis
    (large enough that the rescale matters: factors of about 1 +- 0.02)."""
    g = L.Lattice(*CASE["n"], periodic=CASE["periodic"], tau=CASE["tau"], precision=precision,
                  body_accel=(0.0, 0.0, 0.0), voxel_mode="overlap")
    g.set_periodic_pressure(2, 1.02, 0.98)
    return g


@pytest.mark.parametrize("nsteps,K", [(1, 1), (4, 2), (5, -3)])
def test_adjoint_transpose_identity_periodic_pressure(nsteps, K):
    """<(S(f0 + e u) - S(f0 - e u)) / 2e, w> = <u, S^T w> for nsteps steps of
    collide+stream with the periodic-pressure rescale (its plane means make
    the step nonlinear in f; the central difference is exact to O(e^2))."""
    n = CASE["n"]
    f0 = initial_populations(n)
    x, v, w, r, ids = spheres()
    g = pp_lattice()
    g.set_populations(f0)
    g.set_spheres(x, v, w, r, ids)
    rng = np.random.default_rng(11)
    u = rng.normal(size=f0.shape)
    wt = rng.normal(size=f0.shape)
    g.diff_forward(nsteps, K)
    lam0 = g.diff_backward(g_fT=wt, want_f0=True)["f0"]
    eps = 1e-6
    g.set_populations(f0 + eps * u)
    g.diff_forward(nsteps, K)
    fp = g.get_populations()
    g.set_populations(f0 - eps * u)
    g.diff_forward(nsteps, K)
    fm = g.get_populations()
    lhs = float(np.sum((fp - fm) / (2 * eps) * wt))
    rhs = float(np.sum(u * lam0))
    assert lhs == pytest.approx(rhs, rel=1e-7)
    g.diff_release()


def test_adjoint_periodic_pressure_parameters_fd():
    """Drag gradients w.r.t. v, w, x, tau and sf through a window with the
    periodic-pressure rescale, against central differences of the forward."""
    n = CASE["n"]
    f0 = initial_populations(n)
    x, v, w, r, ids = spheres()
    nsteps = 6
    G = np.random.default_rng(5).normal(size=(nsteps, 3, 6))

    def J(xx, vv, ww, tau):
        g = pp_lattice()
        g.set_tau(tau)
        g.set_populations(f0)
        g.set_spheres(xx, vv, ww, r, ids)
        ft = g.diff_forward(nsteps, 2)
        g.close()
        return float(np.sum(ft * G))

    g = pp_lattice()
    g.set_populations(f0)
    g.set_spheres(x, v, w, r, ids)
    g.diff_forward(nsteps, 2)
    gr = g.diff_backward(G)
    g.close()
    eps = 1e-6
    checks = [("v", (0, 2)), ("v", (2, 0)), ("w", (1, 1)), ("x", (0, 0)), ("x", (2, 2))]
    for name, (p_, d_) in checks:
        arrs = {"x": x.copy(), "v": v.copy(), "w": w.copy()}
        arrs[name][p_, d_] += eps
        jp = J(arrs["x"], arrs["v"], arrs["w"], CASE["tau"])
        arrs[name][p_, d_] -= 2 * eps
        jm = J(arrs["x"], arrs["v"], arrs["w"], CASE["tau"])
        fd = (jp - jm) / (2 * eps)
        assert gr[name][p_, d_] == pytest.approx(fd, rel=1e-5, abs=1e-9), (name, p_, d_)
    fd = (J(x, v, w, CASE["tau"] + eps) - J(x, v, w, CASE["tau"] - eps)) / (2 * eps)
    assert gr["tau"] == pytest.approx(fd, rel=1e-5)


def channel_case(precision, nsteps=40):
    """A reduced differentiable channel: walls x/y, periodic z, body force."""
    n = (32, 32, 48)
    x, v, w, r, ids = L.generate_spheres(n, 6, radius_law=L.RADIUS_CONST, a=5.0, periodic=(0, 0, 1),
                                         gap=1.0, wall_gap=1.5, sigma_v=0.005, sigma_w=0.0005, seed=21)
    lat = L.Lattice(*n, periodic=(0, 0, 1), tau=0.8, precision=precision, body_accel=(0, 0, 1e-6),
                    voxel_mode="overlap")
    lat.init_equilibrium()
    f0 = lat.get_populations()
    return lat, f0, x, v, w, r, ids, nsteps


def test_adjoint_central_differences_20_parameters():
    lat, f0, x, v, w, r, ids, nsteps = channel_case("fp64")
    nc = lat.nc
    G = np.zeros((nsteps, len(r), 6))
    G[:, :, 2] = 1.0    # total z-drag summed over the window
    tx = torch.tensor(x, requires_grad=True)
    tv = torch.tensor(v, requires_grad=True)
    tw = torch.tensor(w, requires_grad=True)
    ttau = torch.tensor(0.8, dtype=torch.float64, requires_grad=True)
    tdsf = torch.zeros(nc, dtype=torch.float64, requires_grad=True)
    ft = psm_window(lat, f0, tx, tv, tw, r, ids, ttau, tdsf, nsteps=nsteps, checkpoint_every=8)
    (ft * torch.tensor(G)).sum().backward()
    rec = lat.cell_records()
    partial = np.nonzero((rec["sf"] > 0.05) & (rec["sf"] < 0.95))[0]
    rng = np.random.default_rng(9)
    params = ([("x", (int(rng.integers(6)), int(rng.integers(3)))) for _ in range(6)]
              + [("v", (int(rng.integers(6)), int(rng.integers(3)))) for _ in range(5)]
              + [("w", (int(rng.integers(6)), int(rng.integers(3)))) for _ in range(4)]
              + [("tau", ())]
              + [("sf", (int(c),)) for c in rng.choice(partial, 4, replace=False)])
    assert len(params) == 20
    grads = {"x": tx.grad.numpy(), "v": tv.grad.numpy(), "w": tw.grad.numpy(),
             "tau": np.array(ttau.grad.item()), "sf": tdsf.grad.numpy()}
    scale = {k: np.abs(grads[k]).max() for k in grads}

    def J(**kw):
        a = dict(x=tx.detach().clone(), v=tv.detach().clone(), w=tw.detach().clone(),
                 tau=ttau.detach().clone(), sf=tdsf.detach().clone())
        for k, (ix, dv) in kw.items():
            a[k][ix] += dv
        with torch.no_grad():
            out = psm_window(lat, f0, a["x"], a["v"], a["w"], r, ids, a["tau"], a["sf"], nsteps=nsteps,
                             checkpoint_every=nsteps)
        return float((out.numpy() * G).sum())

    eps = 1e-6
    for name, ix in params:
        fd = (J(**{name: (ix, eps)}) - J(**{name: (ix, -eps)})) / (2 * eps)
        ad = float(grads[name][ix])
        assert abs(ad - fd) <= 1e-5 * max(abs(fd), 1e-3 * scale[name]), (name, ix, ad, fd)


def test_fp32_adjoint_against_fp64():
    res = {}
    for prec in ("fp64", "fp32_dev"):
        lat, f0, x, v, w, r, ids, nsteps = channel_case(prec)
        lat.set_populations(f0)
        lat.set_spheres(x, v, w, r, ids)
        ft = lat.diff_forward(nsteps, 10)
        G = np.zeros_like(ft)
        G[:, :, 2] = 1.0
        res[prec] = (ft, lat.diff_backward(G))
    ft64, g64 = res["fp64"]
    ft32, g32 = res["fp32_dev"]
    assert rel(ft32[:, :, 2].sum(0), ft64[:, :, 2].sum(0)) < 1e-4
    for k in ("v", "w", "x", "sf"):
        assert rel(g32[k], g64[k]) < 1e-3, k
    assert g32["tau"] == pytest.approx(g64["tau"], rel=1e-3)


def test_window_errors():
    g = gpu_lattice("subcell")
    with pytest.raises(L.PSMError):
        g.diff_forward(2, 1)          # no particles yet
    x, v, w, r, ids = spheres()
    g.set_spheres(x, v, w, r, ids)
    with pytest.raises(L.PSMError):
        g.diff_backward(np.zeros((2, 3, 6)))   # no forward yet
    g.diff_forward(2, 1)
    g.set_tau(0.9)
    with pytest.raises(L.PSMError):
        g.diff_backward(np.zeros((2, 3, 6)))   # tau changed since the forward
    bad = np.zeros(g.nc)
    bad[0] = 0.5                      # cell 0 holds no record
    with pytest.raises(L.PSMError):
        g.perturb_solid_fraction(bad)
    h = L.Lattice(8, 8, 8, precision="fp16_store")
    h.set_spheres(np.array([[4.0, 4.0, 4.0]]), np.zeros((1, 3)), np.zeros((1, 3)), np.array([2.0]),
                  np.array([1]))
    with pytest.raises(L.PSMError):
        h.diff_forward(1, 1)


def test_adjoint_differentiable_channel_full_size():
    """The 128x128x256 differentiable channel (walls x/y, body force 1e-6 along
    z, 200 spheres r=5, v~N(0,0.005)), adjoint of 500 steps with a checkpoint
    every 50: d(total z-drag) against central differences (eps 1e-6, FP64)
    on 20 random parameters, for the FP64 and the FP32-deviation adjoint."""
    n, nsteps, K = (128, 128, 256), 500, 50
    x, v, w, r, ids = L.generate_spheres(n, 200, L.RADIUS_CONST, a=5.0, periodic=(0, 0, 1), gap=0.5,
                                         wall_gap=1.5, sigma_v=0.005, sigma_w=0.0, seed=44)
    grads = {}
    for prec in ("fp64", "fp32_dev"):
        lat = L.Lattice(*n, periodic=(0, 0, 1), tau=0.8, precision=prec, body_accel=(0, 0, 1e-6),
                        voxel_mode="overlap")
        lat.set_spheres(x, v, w, r, ids)
        ft = lat.diff_forward(nsteps, K)
        G = np.zeros_like(ft)
        G[:, :, 2] = 1.0
        grads[prec] = lat.diff_backward(G)
        if prec == "fp64":
            rec = lat.cell_records()
        lat.close()
    lat = L.Lattice(*n, periodic=(0, 0, 1), tau=0.8, precision="fp64", body_accel=(0, 0, 1e-6),
                    voxel_mode="overlap")
    lat.init_equilibrium()
    f0 = lat.get_populations()
    rng = np.random.default_rng(13)
    partial = np.nonzero((rec["sf"] > 0.05) & (rec["sf"] < 0.95))[0]
    params = ([("x", (int(rng.integers(200)), int(rng.integers(3)))) for _ in range(5)]
              + [("v", (int(rng.integers(200)), int(rng.integers(3)))) for _ in range(6)]
              + [("w", (int(rng.integers(200)), int(rng.integers(3)))) for _ in range(4)]
              + [("tau", ())]
              + [("sf", (int(c),)) for c in rng.choice(partial, 4, replace=False)])
    assert len(params) == 20

    def J(name, ix, dv):
        a = {"x": x.copy(), "v": v.copy(), "w": w.copy(), "tau": np.array(0.8), "sf": np.zeros(lat.nc)}
        a[name][ix] += dv
        lat.set_tau(float(a["tau"]))
        lat.set_populations(f0)
        lat.clear_records()
        lat.set_spheres(a["x"], a["v"], a["w"], r, ids)
        if name == "sf":
            lat.perturb_solid_fraction(a["sf"])
        return float(lat.diff_forward(nsteps, nsteps)[:, :, 2].sum())

    eps = 1e-6
    for name, ix in params:
        fd = (J(name, ix, eps) - J(name, ix, -eps)) / (2 * eps)
        g64 = float(np.asarray(grads["fp64"][name])[ix])
        g32 = float(np.asarray(grads["fp32_dev"][name])[ix])
        scale = np.abs(grads["fp64"][name]).max()
        assert abs(g64 - fd) <= 1e-5 * max(abs(fd), 1e-3 * scale), (name, ix, g64, fd)
        assert abs(g32 - fd) <= 2e-3 * max(abs(fd), 1e-2 * scale), (name, ix, g32, fd)
