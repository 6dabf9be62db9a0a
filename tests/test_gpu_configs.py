"""The benchmark configurations against the oracle, through the C ABI.

- Pressure-driven channel (the 512x512x768 configuration at a size the
  oracle runs in seconds): walls in x/y, Zhang-Kwok periodic pressure along
  z from a Poiseuille initial field, tau 0.55, radii U[3,6], slow prescribed
  kinematics.  FP64: records bit-exact every step, populations, forces and
  torques within 1e-12.  FP32 deviation storage: per-particle drag within
  1e-5 of the oracle's.
- Polydisperse suspension (the 384^3 configuration, reduced): radii
  lognormal(5, 0.25) clipped to [3, 10], tau 0.6, FP32 deviation and FP16
  storage against the oracle.
- Dense suspension at its own size, 256^3 with 2000 spheres r~U[4,8],
  tau 0.7: coupled steps with bit-exact sub-cell records and populations
  and forces within 1e-12.
- The channel at its own size, 512x512x768 with 100000 spheres: FP32
  deviation storage against FP64 on the device, the drag of 1000 particles
  checked per particle.

Per-particle errors are |dF_p| / max(|F_p|, floor), the floor a stated
fraction of the median |F| so that a particle with nearly zero load does
not divide by round-off.
"""
import os

import numpy as np
import pytest

import lbdem_psm as L
from oracle import oracle as O

pytestmark = pytest.mark.gpu

THREADS = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def rel_err(a, b):
    den = np.abs(b).max()
    return float(np.abs(a - b).max() / den) if den > 0 else float(np.abs(a - b).max())


def torque_err(Tg, To, Fo, arm=10.0):
    den = max(np.abs(To).max(), np.abs(Fo).max() * arm)
    return float(np.abs(Tg - To).max() / den) if den > 0 else float(np.abs(Tg - To).max())


def particle_rel(a, ref, floor_frac=1e-2):
    mag = np.linalg.norm(ref, axis=1)
    floor = floor_frac * float(np.median(mag))
    return np.linalg.norm(a - ref, axis=1) / np.maximum(mag, floor)


def channel_spheres(n, count, seed):
    return L.generate_spheres(n, count, L.RADIUS_UNIFORM, a=3.0, b=6.0, periodic=(0, 0, 1), gap=0.0,
                              wall_gap=2.0, sigma_v=0.002, sigma_w=2e-4, seed=seed)


def setup_channel(lat, n, drop):
    lat.init_poiseuille(2, 1.0, 1.0 - drop, 0.05)
    lat.set_periodic_pressure(2, 1.0, 1.0 - drop)


def test_pressure_channel_config_fp64_matches_oracle():
    n = (64, 64, 96)
    drop = 1e-4 * 96 / 768          # the large channel's pressure gradient
    x, v, w, r, ids = channel_spheres(n, 150, 31)
    g = L.Lattice(*n, periodic=(0, 0, 1), tau=0.55, precision="fp64", slices_per_dim=8)
    o = O.OracleLattice(*n, periodic=(0, 0, 1), tau=0.55, slices_per_dim=8)
    o.set_threads(THREADS)
    try:
        setup_channel(g, n, drop)
        setup_channel(o, n, drop)
        worst = 0.0
        for s in range(8):
            if s:
                x = L.advance_positions(x, v, n, (0, 0, 1))
            g.set_spheres(x, v, w, r, ids)
            o.set_spheres(x, v, w, r, ids)
            rg, ro = g.cell_records(), o.cell_records()
            assert np.array_equal(rg["sf"], ro["sf"]) and np.array_equal(rg["id"], ro["id"])
            assert np.array_equal(rg["us"], ro["us"])
            g.collide_stream(1)
            o.collide_stream(1, THREADS)
            o.set_threads(1)
            Fo, To = o.sum_force_torque(x)
            o.set_threads(THREADS)
            Fg, Tg = g.sum_force_torque(x)
            worst = max(worst, rel_err(Fg, Fo), torque_err(Tg, To, Fo))
        assert worst < 1e-12, worst
        assert rel_err(g.get_populations(), o.get_populations()) < 1e-12
        assert g.stats()["records"] > 0.15 * np.prod(n)
    finally:
        g.close()
        o.close()


def _precision_vs_oracle(n, periodic, tau, spheres, steps, precision, setup=None):
    x, v, w, r, ids = spheres
    g = L.Lattice(*n, periodic=periodic, tau=tau, precision=precision, slices_per_dim=8)
    o = O.OracleLattice(*n, periodic=periodic, tau=tau, slices_per_dim=8)
    o.set_threads(THREADS)
    try:
        if setup:
            setup(g)
            setup(o)
        for s in range(steps):
            if s:
                x = L.advance_positions(x, v, n, periodic)
            g.set_spheres(x, v, w, r, ids)
            o.set_spheres(x, v, w, r, ids)
            g.collide_stream(1)
            o.collide_stream(1, THREADS)
        o.set_threads(1)
        Fo, To = o.sum_force_torque(x)
        Fg, Tg = g.sum_force_torque(x)
        _, uo = L.macroscopic(o.get_populations())
        _, ug = L.macroscopic(g.get_populations())
    finally:
        g.close()
        o.close()
    return Fg, Tg, Fo, To, float(np.linalg.norm(ug - uo) / np.linalg.norm(uo))


def test_pressure_channel_config_fp32_drag_per_particle():
    n = (64, 64, 96)
    drop = 1e-4 * 96 / 768
    sph = channel_spheres(n, 150, 31)
    Fg, Tg, Fo, To, uerr = _precision_vs_oracle(n, (0, 0, 1), 0.55, sph, 20, "fp32_dev",
                                                setup=lambda lat: setup_channel(lat, n, drop))
    e = particle_rel(Fg, Fo)
    et = particle_rel(Tg, To)
    print("fp32 channel: drag per particle median %.2e p99 %.2e max %.2e; norm-wise %.2e; torque max %.2e; "
          "velocity L2 %.2e" % (np.median(e), np.percentile(e, 99), e.max(), rel_err(Fg, Fo), et.max(), uerr))
    assert rel_err(Fg, Fo) < 1e-5
    assert e.max() < 1e-5
    assert uerr < 1e-5


@pytest.mark.parametrize("precision,drag_tol,particle_tol", [("fp32_dev", 1e-5, 1e-5), ("fp16_store", 1e-3, 2e-3)])
def test_polydisperse_config_precision_vs_oracle(precision, drag_tol, particle_tol):
    n = (72, 72, 72)
    sph = L.generate_spheres(n, 130, L.RADIUS_LOGNORMAL, a=5.0, b=0.25, lo=3.0, hi=10.0, periodic=(1, 1, 1),
                             gap=0.0, sigma_v=0.001, sigma_w=1e-4, seed=41)
    Fg, Tg, Fo, To, uerr = _precision_vs_oracle(n, (1, 1, 1), 0.6, sph, 20, precision)
    e = particle_rel(Fg, Fo)
    print("%s polydisperse: drag per particle median %.2e p99 %.2e max %.2e; norm-wise %.2e; torque %.2e; "
          "velocity L2 %.2e" % (precision, np.median(e), np.percentile(e, 99), e.max(), rel_err(Fg, Fo),
                                rel_err(Tg, To), uerr))
    assert rel_err(Fg, Fo) < drag_tol
    assert e.max() < particle_tol


def test_dense_suspension_config_full_size_matches_oracle():
    n = (256, 256, 256)
    x, v, w, r, ids = L.generate_spheres(n, 2000, L.RADIUS_UNIFORM, a=4.0, b=8.0, periodic=(1, 1, 1), gap=0.0,
                                         sigma_v=0.01, sigma_w=1e-3, seed=2024)
    g = L.Lattice(*n, periodic=(1, 1, 1), tau=0.7, precision="fp64", slices_per_dim=8)
    o = O.OracleLattice(*n, periodic=(1, 1, 1), tau=0.7, slices_per_dim=8)
    o.set_threads(THREADS)
    try:
        for s in range(2):
            if s:
                x = L.advance_positions(x, v, n, (1, 1, 1))
            g.set_spheres(x, v, w, r, ids)
            o.set_spheres(x, v, w, r, ids)
            rg, ro = g.cell_records(), o.cell_records()
            assert np.array_equal(rg["sf"], ro["sf"]) and np.array_equal(rg["id"], ro["id"])
            assert np.array_equal(rg["us"], ro["us"])
            del rg, ro
            g.collide_stream(1)
            o.collide_stream(1, THREADS)
            o.set_threads(1)
            Fo, To = o.sum_force_torque(x)
            o.set_threads(THREADS)
            Fg, Tg = g.sum_force_torque(x)
            assert rel_err(Fg, Fo) < 1e-12 and torque_err(Tg, To, Fo) < 1e-12
        assert np.array_equal(g.get_populations(), o.get_populations())
        st = g.stats()
        assert st["records"] > 2_000_000 and st["images"] > 0
    finally:
        g.close()
        o.close()


def test_channel_config_full_size_fp32_spot_check():
    n = (512, 512, 768)
    x, v, w, r, ids = channel_spheres(n, 100000, 2024)
    out = {}
    for prec in ("fp64", "fp32_dev"):
        g = L.Lattice(*n, periodic=(0, 0, 1), tau=0.55, precision=prec, slices_per_dim=8)
        try:
            setup_channel(g, n, 1e-4)
            g.set_spheres(x, v, w, r, ids)
            g.step(10, advance=True)
            out[prec] = g.particle_forces()
            rc, umax, nf = g.check_health()
            assert rc == 0 and nf == 0
        finally:
            g.close()
    sub = np.random.default_rng(7).choice(len(r), 1000, replace=False)
    e = particle_rel(out["fp32_dev"][0][sub], out["fp64"][0][sub])
    print("channel 512x512x768 fp32 vs fp64 drag, 1000 particles: median %.2e p99 %.2e max %.2e"
          % (np.median(e), np.percentile(e, 99), e.max()))
    assert e.max() < 1e-5
    assert rel_err(out["fp32_dev"][0], out["fp64"][0]) < 1e-5
