"""GPU (HIP, through the C ABI) versus the CPU oracle on the same seeded inputs.

Bars: FP64 populations, per-cell records (solid fraction, particle id,
particle velocity, per-cell force) and sub-cell counts bit-exact (the FP64
kernels evaluate the oracle's expressions in the same order without
contraction); per-particle force and torque within 1e-12 relative error
(only the summation order differs: a parallel tree here, a lexicographic
scan in the reference).
"""
import numpy as np
import pytest

import lbdem_psm as L
from oracle import oracle as O

pytestmark = pytest.mark.gpu

FORCE_RTOL = 1e-12


def rel_err(a, b):
    den = np.abs(b).max()
    return float(np.abs(a - b).max() / den) if den > 0 else float(np.abs(a - b).max())


def torque_err(Tg, To, Fo, arm=10.0):
    """Torque error on the scale of the torques a force of the run's size can
    make: |dT| / max(|T|, |F| * arm) (a symmetric case has round-off torques)."""
    den = max(np.abs(To).max(), np.abs(Fo).max() * arm)
    return float(np.abs(Tg - To).max() / den) if den > 0 else float(np.abs(Tg - To).max())


def assert_records_equal(rg, ro):
    assert np.array_equal(rg["sf"], ro["sf"])
    assert np.array_equal(rg["id"], ro["id"])
    assert np.array_equal(rg["us"], ro["us"])


def make_pair(n, periodic=(1, 1, 1), tau=0.7, N=5, body=(0.0, 0.0, 0.0), precision="fp64"):
    g = L.Lattice(*n, periodic=periodic, tau=tau, precision=precision, slices_per_dim=N,
                  body_accel=body)
    o = O.OracleLattice(*n, periodic=periodic, tau=tau, slices_per_dim=N, body_accel=body)
    return g, o


def random_state(n, seed, u0=0.02):
    rng = np.random.default_rng(seed)
    nc = int(np.prod(n))
    rho = 1 + rng.normal(0, 0.01, nc)
    u = rng.normal(0, u0, (3, nc))
    return rho, u


@pytest.mark.parametrize("N", [5, 8, 6])
def test_subcell_counts_bitexact(N):
    rng = np.random.default_rng(10 + N)
    m = 20000
    r = rng.uniform(2.5, 10.0, m)
    d = rng.normal(0, 1, (m, 3))
    d *= ((r + rng.uniform(-1.3, 1.3, m)) / np.linalg.norm(d, axis=1))[:, None]
    # exact grid offsets and dyadic centres too
    d[:2000] = np.round(d[:2000] * 8) / 8
    with L.Lattice(4, 4, 4, slices_per_dim=N) as g:
        sf, cnt, br = g.subcell_counts(d, r)
    for k in range(m):
        s, n_, b = O.calc_solid_fraction(*d[k], r[k], N)
        assert (sf[k], cnt[k], br[k]) == (s, n_, b), k


@pytest.mark.parametrize("periodic", [(1, 1, 1), (0, 1, 0), (0, 0, 1)])
def test_fluid_populations_bitexact(periodic):
    n = (12, 10, 14)
    g, o = make_pair(n, periodic, tau=0.63)
    rho, u = random_state(n, 1)
    g.init_equilibrium(rho, u)
    o.init_equilibrium(rho, u)
    f0 = o.get_populations()
    assert np.array_equal(g.get_populations(), f0)
    g.collide_stream(10)
    o.collide_stream(10)
    assert np.array_equal(g.get_populations(), o.get_populations())
    mg, jg = g.conservation()
    mo, jo = o.conservation()
    assert abs(mg - mo) < 1e-13 and np.allclose(jg, jo, atol=1e-13)
    g.close()


def test_guo_forcing_with_walls_bitexact():
    n = (9, 6, 8)
    g, o = make_pair(n, (0, 1, 1), tau=0.9, body=(0.0, 2e-6, 1e-5))
    rho, u = random_state(n, 2, 0.01)
    g.init_equilibrium(rho, u)
    o.init_equilibrium(rho, u)
    g.collide_stream(20)
    o.collide_stream(20)
    assert np.array_equal(g.get_populations(), o.get_populations())
    g.close()


def spheres(n, count, seed, rlo=2.5, rhi=4.5, periodic=(1, 1, 1)):
    return L.generate_spheres(n, count, L.RADIUS_UNIFORM, a=rlo, b=rhi, periodic=periodic,
                              gap=0.0, wall_gap=1.0, sigma_v=0.01, sigma_w=1e-3, seed=seed)


def run_coupled(n, periodic, N, count, steps, seed, tau=0.7, init_vel=False, body=(0, 0, 0),
                place_on_faces=True):
    g, o = make_pair(n, periodic, tau=tau, N=N, body=body)
    x, v, w, r, ids = spheres(n, count, seed, periodic=periodic)
    if place_on_faces:
        # push a few spheres across the periodic faces to exercise images
        for k, d in enumerate([0, 1, 2]):
            if periodic[d] and k < count:
                x[k, d] = -0.3 if k % 2 == 0 else n[d] - 0.6
    rho, u = random_state(n, seed, 0.01)
    g.init_equilibrium(rho, u)
    o.init_equilibrium(rho, u)
    worst = 0.0
    for s in range(steps):
        if s > 0:
            x = L.advance_positions(x, v, n, periodic)
        o.set_spheres(x, v, w, r, ids, init_vel=(init_vel and s == 0))
        g.set_spheres(x, v, w, r, ids, init_vel=(init_vel and s == 0))
        assert_records_equal(g.cell_records(), o.cell_records())
        if init_vel and s == 0:
            assert np.array_equal(g.get_populations(), o.get_populations())
        o.collide_stream(1)
        g.collide_stream(1)
        rg, ro = g.cell_records(), o.cell_records()
        assert np.array_equal(rg["F"], ro["F"])
        Fo, To = o.sum_force_torque(x)
        Fg, Tg = g.sum_force_torque(x)
        worst = max(worst, rel_err(Fg, Fo), torque_err(Tg, To, Fo))
        assert rel_err(Fg, Fo) < FORCE_RTOL and torque_err(Tg, To, Fo) < FORCE_RTOL
    assert np.array_equal(g.get_populations(), o.get_populations())
    st = g.stats()
    g.close()
    return worst, st


@pytest.mark.parametrize("N", [5, 8, 6])
def test_coupled_steps_periodic(N):
    worst, st = run_coupled((32, 28, 30), (1, 1, 1), N, 14, 4, seed=3 + N)
    assert st["images"] > 0 and st["records"] > 0
    print("worst force rel err", worst, st)


def test_coupled_steps_walls_and_force():
    run_coupled((30, 26, 32), (0, 0, 1), 8, 10, 4, seed=21, body=(0.0, 0.0, 1e-6))


def test_coupled_init_vel():
    run_coupled((24, 24, 24), (1, 1, 1), 5, 6, 2, seed=4, init_vel=True)


def test_ownership_history_parity():
    n = (16, 16, 16)
    for order in ((0, 1), (1, 0)):
        for prior in (None, (0.9, 2), (0.2, 1)):
            g, o = make_pair(n, (0, 0, 0), tau=0.8, N=5)
            xa = np.array([8.0 - 2.9, 8.0, 8.0])
            xb = np.array([8.0 + 3.15, 8.0, 8.0])
            X = np.array([xa, xb])[list(order)]
            ids = np.array([1, 2])[list(order)]
            if prior is not None:
                sf = np.zeros(16 ** 3)
                pid = np.zeros(16 ** 3, dtype=np.int64)
                c = (8 * 16 + 8) * 16 + 8
                sf[c], pid[c] = prior
                us = np.zeros((3, 16 ** 3))
                us[:, c] = [0.01, 0.02, 0.03]
                g.set_cell_records(sf, pid, us)
                o.set_cell_records(sf, pid, us)
            V = np.array([[0.01, 0, 0], [-0.01, 0.0, 0.0]])
            g.set_spheres(X, V, np.zeros((2, 3)), [3.0, 3.0], ids, auto_images=False)
            o.set_spheres(X, V, np.zeros((2, 3)), [3.0, 3.0], ids, auto_images=False)
            assert_records_equal(g.cell_records(), o.cell_records())
            g.collide_stream(1)
            o.collide_stream(1)
            Fg, Tg = g.sum_force_torque(X)
            Fo, To = o.sum_force_torque(X)
            assert rel_err(Fg, Fo) < FORCE_RTOL and torque_err(Tg, To, Fo) < FORCE_RTOL
            g.close()


def test_stale_records_are_kept_and_summed():
    """A record outside every bounding box keeps its values (the reference only
    visits cells inside boxes) and still contributes its force to its owner."""
    n = (16, 16, 16)
    g, o = make_pair(n, (1, 1, 1), tau=0.8, N=5)
    sf = np.zeros(16 ** 3)
    pid = np.zeros(16 ** 3, dtype=np.int64)
    us = np.zeros((3, 16 ** 3))
    c = (1 * 16 + 1) * 16 + 1
    sf[c], pid[c] = 0.5, 1
    us[:, c] = [0.02, 0.0, -0.01]
    for lat in (g, o):
        lat.set_cell_records(sf, pid, us)
    x = np.array([[9.0, 9.0, 9.0]])
    for lat in (g, o):
        lat.set_spheres(x, np.zeros((1, 3)), np.zeros((1, 3)), [2.0], [1])
    assert_records_equal(g.cell_records(), o.cell_records())
    assert g.cell_records()["sf"][c] == 0.5
    g.collide_stream(1)
    o.collide_stream(1)
    Fg, Tg = g.sum_force_torque(x)
    Fo, To = o.sum_force_torque(x)
    assert g.stats()["orphans"] >= 1
    assert rel_err(Fg, Fo) < FORCE_RTOL and torque_err(Tg, To, Fo) < FORCE_RTOL
    g.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32_dev"])
def test_stale_record_sums_are_deterministic(precision):
    """Stale records of several particles -- every record after
    set_cell_records, and the records a particle leaves behind when it jumps
    beyond its bounding box -- are summed per owner in cell order: two
    handles give bit-identical forces and torques, and FP64 matches the
    oracle within 1e-12."""
    n = (40, 36, 32)
    x, v, w, r, ids = spheres(n, 12, 91, 3.0, 5.0)
    ref = O.OracleLattice(*n, tau=0.7, slices_per_dim=8)
    ref.set_spheres(x + 0.4, v, w, r, ids)
    prior = ref.cell_records()
    ref.close()
    jump = x.copy()
    jump[::2, 2] = (jump[::2, 2] + 9.0) % n[2]        # half of them leave their records behind
    runs = []
    for _ in range(2):
        g = L.Lattice(*n, tau=0.7, precision=precision, slices_per_dim=8)
        o = O.OracleLattice(*n, tau=0.7, slices_per_dim=8) if precision == "fp64" else None
        out = []
        for lat in ([g, o] if o else [g]):
            lat.set_spheres(x, v, w, r, ids)
            lat.set_cell_records(prior["sf"], prior["id"], prior["us"])   # all stale
            lat.collide_stream(1)
            F1, T1 = lat.sum_force_torque(x)
            lat.set_spheres(jump, v, w, r, ids)                          # records left behind
            lat.collide_stream(1)
            F2, T2 = lat.sum_force_torque(jump)
            out.append((F1, T1, F2, T2))
        assert g.stats()["orphans"] > 0
        if o:
            for a, b in ((0, 1), (2, 3)):
                assert rel_err(out[0][a], out[1][a]) < FORCE_RTOL
                assert torque_err(out[0][b], out[1][b], out[1][a]) < FORCE_RTOL
            o.close()
        runs.append(out[0])
        g.close()
    for a, b in zip(runs[0], runs[1]):
        assert np.array_equal(a, b)
    assert np.abs(runs[0][0]).max() > 1e-6 and np.abs(runs[0][2]).max() > 1e-6


def test_many_candidates_in_one_bin():
    """More than 64 small spheres whose boxes meet one 8^3 bin: the force sum
    takes the positions of entries past the 64th from memory, and the fold
    walks a long candidate list.  Overlapping spheres go through the
    four-case rule like any other."""
    n = (24, 24, 24)
    g, o = make_pair(n, (1, 1, 1), tau=0.7, N=8)
    rng = np.random.default_rng(77)
    m = 90
    x = rng.uniform(7.6, 15.4, (m, 3))
    v = rng.normal(0, 0.01, (m, 3))
    w = rng.normal(0, 1e-3, (m, 3))
    r = rng.uniform(0.6, 1.1, m)
    # every sphere's box (r + sqrt(3.1)/2 around the centre) meets the bin of cells 8..15
    assert np.all(x - r[:, None] - 0.89 < 16) and np.all(x + r[:, None] + 0.89 >= 8)
    ids = list(range(1, m + 1))
    rho, u = random_state(n, 9, 0.01)
    for lat in (g, o):
        lat.init_equilibrium(rho, u)
    for s in range(2):
        if s > 0:
            x = L.advance_positions(x, v, n, (1, 1, 1))
        for lat in (g, o):
            lat.set_spheres(x, v, w, r, ids)
        assert_records_equal(g.cell_records(), o.cell_records())
        g.collide_stream(1)
        o.collide_stream(1)
        Fg, Tg = g.sum_force_torque(x)
        Fo, To = o.sum_force_torque(x)
        assert rel_err(Fg, Fo) < FORCE_RTOL and torque_err(Tg, To, Fo) < FORCE_RTOL
    g.close()


def test_user_ghosts_fold_into_owner():
    """Ghost copies with the owner's tag (LIGGGHTS-style, no automatic images):
    forces land on map(tag) = the first index, ghost slots stay zero."""
    n = (20, 20, 20)
    g, o = make_pair(n, (1, 1, 1), tau=0.7, N=5)
    x = np.array([[0.4, 10.0, 10.0], [8.0, 8.0, 3.0], [20.4, 10.0, 10.0]])   # 2 = ghost of 0
    v = np.array([[0.01, 0, 0], [0, -0.01, 0.005], [0.01, 0, 0]])
    w = np.zeros((3, 3))
    ids = [5, 6, 5]
    rho, u = random_state(n, 8, 0.01)
    for lat in (g, o):
        lat.init_equilibrium(rho, u)
        lat.set_spheres(x, v, w, [3.0, 2.5, 3.0], ids, auto_images=False)
    assert_records_equal(g.cell_records(), o.cell_records())
    g.collide_stream(1)
    o.collide_stream(1)
    Fg, Tg = g.sum_force_torque(x)
    Fo, To = o.sum_force_torque(x)
    assert np.all(Fg[2] == 0) and np.all(Tg[2] == 0)
    assert rel_err(Fg, Fo) < FORCE_RTOL and torque_err(Tg, To, Fo) < FORCE_RTOL
    g.close()


def test_exclude_types_parity():
    n = (20, 20, 20)
    g, o = make_pair(n, (1, 1, 1), tau=0.7, N=8)
    x, v, w, r, ids = spheres(n, 6, 12)
    types = [1, 2, 1, 3, 2, 1]
    for lat in (g, o):
        lat.set_spheres(x, v, w, r, ids, types=types, exclude=[2])
    assert_records_equal(g.cell_records(), o.cell_records())
    g.close()


def test_resident_step_matches_host_loop():
    n = (32, 32, 32)
    x, v, w, r, ids = spheres(n, 12, 31)
    a = L.Lattice(*n, tau=0.7, slices_per_dim=8)
    b = L.Lattice(*n, tau=0.7, slices_per_dim=8)
    a.set_spheres(x, v, w, r, ids)
    b.set_spheres(x, v, w, r, ids)
    # a: one host-driven step then resident steps; b: host loop throughout
    a.collide_stream(1)
    a.sum_force_torque(x)
    a.step(3, advance=True)
    b.collide_stream(1)
    b.sum_force_torque(x)
    xx = x
    for _ in range(3):
        xx = L.advance_positions(xx, v, n, (1, 1, 1))
        b.set_spheres(xx, v, w, r, ids)
        b.collide_stream(1)
        Fb, Tb = b.sum_force_torque(xx)
    Fa, Ta = a.particle_forces()
    assert np.array_equal(a.particles(), xx)
    assert np.array_equal(a.get_populations(), b.get_populations())
    assert np.array_equal(Fa, Fb) and np.array_equal(Ta, Tb)
    a.close()
    b.close()


def test_force_sum_is_deterministic():
    n = (40, 40, 40)
    x, v, w, r, ids = spheres(n, 30, 77, 3.0, 6.0)
    out = []
    for _ in range(2):
        g = L.Lattice(*n, tau=0.7, slices_per_dim=8)
        g.set_spheres(x, v, w, r, ids)
        g.step(3, advance=True)
        out.append(g.particle_forces())
        g.close()
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def particle_rel(a, ref, floor_frac=1e-2):
    """|a_p - ref_p| / max(|ref_p|, floor) per particle, the floor a fraction
    of the median |ref| (a particle with nearly zero load would divide by
    round-off)."""
    mag = np.linalg.norm(ref, axis=1)
    return np.linalg.norm(a - ref, axis=1) / np.maximum(mag, floor_frac * float(np.median(mag)))


@pytest.mark.parametrize("precision,drag_tol,particle_tol", [("fp32_dev", 1e-5, 1e-5), ("fp16_store", 1e-3, 2e-3)])
def test_precision_variants_track_fp64(precision, drag_tol, particle_tol):
    n = (48, 48, 48)
    x, v, w, r, ids = spheres(n, 20, 5, 3.0, 6.0)
    res = {}
    for prec in ("fp64", precision):
        g = L.Lattice(*n, tau=0.7, precision=prec, slices_per_dim=8)
        g.set_spheres(x, v, w, r, ids)
        g.step(20, advance=True)
        F, T = g.particle_forces()
        res[prec] = (F, T, g.get_populations())
        g.close()
    F64, T64, f64 = res["fp64"]
    F, T, f = res[precision]
    drag = rel_err(F, F64)
    torque = rel_err(T, T64)
    _, u64 = L.macroscopic(f64)
    _, u = L.macroscopic(f)
    uerr = float(np.abs(u - u64).max() / np.abs(u64).max())
    e = particle_rel(F, F64)
    print(precision, "drag", drag, "per particle max", e.max(), "torque", torque, "velocity", uerr)
    assert drag < drag_tol
    assert e.max() < particle_tol


def test_numeric_health_and_errors():
    g = L.Lattice(8, 8, 8, tau=0.7)
    rc, mu, nf = g.check_health()
    assert rc == 0 and mu == 0.0 and nf == 0
    f = g.get_populations()
    f[3, 17] = np.nan
    g.set_populations(f)
    rc, mu, nf = g.check_health()
    assert rc == L.ERR_NUMERIC and nf == 1
    with pytest.raises(L.PSMError) as e:
        g.set_spheres([[1.0, 1.0, 1.0]], [[0.0, 0, 0]], [[0.0, 0, 0]], [-1.0], [1])
    assert e.value.code == L.ERR_INVALID_ARG
    with pytest.raises(L.PSMError) as e:
        g.sum_force_torque()
    assert e.value.code == L.ERR_STATE
    g.close()


def test_record_capacity_error():
    g = L.Lattice(24, 24, 24, tau=0.7, max_solid_cells=50)
    with pytest.raises(L.PSMError) as e:
        g.set_spheres([[12.0, 12.0, 12.0]], [[0.0, 0, 0]], [[0.0, 0, 0]], [5.0], [1])
    assert e.value.code == L.ERR_CAPACITY
    g.close()


def test_poiseuille_init_matches_oracle():
    n = (12, 10, 14)
    for prec, tol in (("fp64", 0.0), ("fp32_dev", 1e-7)):
        g = L.Lattice(*n, periodic=(1, 0, 0), tau=0.8, precision=prec)
        o = O.OracleLattice(*n, periodic=(1, 0, 0), tau=0.8)
        for dim in (0, 1, 2):
            n3 = list(n)
            n3[dim] -= 1
            g.init_poiseuille(dim, 1.002, 0.998, 0.03, n3=n3)
            o.init_poiseuille(dim, 1.002, 0.998, 0.03, n3=n3)
            fg, fo = g.get_populations(), o.get_populations()
            if tol == 0.0:
                assert np.array_equal(fg, fo)
            else:
                assert np.abs(fg - fo).max() < tol


def test_periodic_pressure_channel_matches_oracle():
    """Pressure-driven channel (periodic x with the Zhang-Kwok boundary,
    walls y/z) with spheres: populations and forces against the oracle."""
    n = (24, 18, 20)
    g, o = make_pair(n, periodic=(1, 0, 0), tau=0.6, N=5)
    for lat in (g, o):
        lat.init_poiseuille(0, 1.001, 0.999, 0.02, n3=(n[0] - 1, n[1], n[2]))
        lat.set_periodic_pressure(0, 1.001, 0.999)
    x = np.array([[8.2, 9.1, 10.3], [17.6, 7.4, 9.8]])
    v = np.array([[0.004, 0.0, 0.001], [-0.002, 0.003, 0.0]])
    w = np.zeros((2, 3))
    r = np.array([3.5, 3.0])
    ids = np.array([1, 2])
    g.set_spheres(x, v, w, r, ids)
    o.set_spheres(x, v, w, r, ids)
    for _ in range(12):
        g.collide_stream(1)
        o.collide_stream(1, 1)
    Fg, Tg = g.sum_force_torque(x)
    Fo, To = o.sum_force_torque(x)
    assert rel_err(g.get_populations(), o.get_populations()) < 1e-12
    assert rel_err(Fg, Fo) < 1e-12
    assert torque_err(Tg, To, Fo) < 1e-12
    # switching it off stops the rescale
    g.set_periodic_pressure(0, 1.0, 1.0, enable=False)
    o.set_periodic_pressure(0, 1.0, 1.0, enable=False)
    g.collide_stream(2)
    o.collide_stream(2, 1)
    assert rel_err(g.get_populations(), o.get_populations()) < 1e-12


@pytest.mark.parametrize("precision,pop_tol,force_tol", [("fp32_dev", 1e-5, 2e-5),
                                                         ("fp16_store", 5e-3, 2e-3)])
@pytest.mark.parametrize("driver", ["body", "pressure"])
def test_walled_channel_precisions_track_oracle(precision, pop_tol, force_tol, driver):
    """The reduced-precision lattices with walls on y and z (half-way
    bounce-back) and moving spheres against the FP64 oracle: driven by a body
    force, or by the Zhang-Kwok periodic pressure on x. Populations and
    velocity are compared on the scale of their largest deviation from rest.
    Measured: FP32 within 2e-6 (populations, velocity) and 4e-6 (forces); FP16
    storage within 1.2e-3 and 1.4e-4."""
    n = (24, 18, 20)
    body = (2e-6, 0.0, 0.0) if driver == "body" else (0.0, 0.0, 0.0)
    g, o = make_pair(n, periodic=(1, 0, 0), tau=0.6, N=5, body=body, precision=precision)
    if driver == "pressure":
        for lat in (g, o):
            lat.init_poiseuille(0, 1.001, 0.999, 0.02, n3=(n[0] - 1, n[1], n[2]))
            lat.set_periodic_pressure(0, 1.001, 0.999)
    x = np.array([[8.2, 9.1, 10.3], [17.6, 7.4, 9.8], [3.1, 13.0, 4.2]])
    v = np.array([[0.004, 0.0, 0.001], [-0.002, 0.003, 0.0], [0.0, -0.001, 0.002]])
    w = np.array([[0.0, 0.0, 0.001], [0.0, 0.0, 0.0], [0.002, 0.0, 0.0]])
    r = np.array([3.5, 3.0, 2.6])
    ids = np.array([1, 2, 3])
    g.set_spheres(x, v, w, r, ids)
    o.set_spheres(x, v, w, r, ids)
    g.collide_stream(16)
    o.collide_stream(16, 1)
    Fg, Tg = g.sum_force_torque(x)
    Fo, To = o.sum_force_torque(x)
    fg, fo = g.get_populations(), o.get_populations()
    _, ug = L.macroscopic(fg)
    _, uo = L.macroscopic(fo)
    print(precision, driver, "pop", rel_err(fg, fo), "u", rel_err(ug, uo),
          "F", rel_err(Fg, Fo), "T", torque_err(Tg, To, Fo))
    assert rel_err(fg, fo) < pop_tol
    assert rel_err(ug, uo) < pop_tol
    assert rel_err(Fg, Fo) < force_tol
    assert torque_err(Tg, To, Fo) < force_tol
    g.close()


# cross-section boxes (a0, a1, b0, b1) over the two other axes in increasing
# order, touching a lattice edge on one side and not the other
BOXES = {0: (0, 9, 3, 15), 1: (2, 13, 0, 11), 2: (0, 13, 1, 6)}


@pytest.mark.parametrize("box", [False, True])
@pytest.mark.parametrize("dim", [0, 1, 2])
def test_periodic_pressure_every_axis_matches_oracle(dim, box):
    """The Zhang-Kwok rescale along each axis of a fully periodic box with
    spheres, 6 steps, against the oracle (the inlet/outlet boxes are x, y
    or z planes; each axis maps the box onto the lattice differently), with
    the default interior cross-section and with an uneven box.  The device
    takes the plane sums from the collide's own loads of the plane cells
    inside the box."""
    n = (14, 12, 16)
    g, o = make_pair(n, periodic=(1, 1, 1), tau=0.8, N=5)
    for lat in (g, o):
        lat.set_periodic_pressure(dim, 1.02, 0.98, range4=BOXES[dim] if box else None)
    x = np.array([[6.3, 5.2, 3.1], [7.6, 4.1, 10.4], [4.2, 8.7, 14.6]])
    v = np.array([[0.01, -0.004, 0.003], [-0.006, 0.002, 0.008], [0.003, 0.005, -0.007]])
    w = np.zeros((3, 3))
    r = np.array([3.1, 2.7, 2.4])
    ids = np.array([1, 2, 3])
    g.set_spheres(x, v, w, r, ids)
    o.set_spheres(x, v, w, r, ids)
    for _ in range(6):
        g.collide_stream(1)
        o.collide_stream(1, 1)
    Fg, Tg = g.sum_force_torque(x)
    Fo, To = o.sum_force_torque(x)
    assert rel_err(g.get_populations(), o.get_populations()) < 1e-12
    assert rel_err(Fg, Fo) < 1e-12
    assert torque_err(Tg, To, Fo) < 1e-12


def test_periodic_pressure_errors():
    g = L.Lattice(8, 8, 8, periodic=(1, 0, 1), tau=0.8)
    with pytest.raises(L.PSMError):
        g.set_periodic_pressure(1, 1.01, 0.99)          # y is not periodic
    with pytest.raises(L.PSMError):
        g.set_periodic_pressure(0, 1.01, 0.99, range4=(0, 8, 0, 7))   # box beyond the lattice
    with pytest.raises(L.PSMError):
        g.set_periodic_pressure(0, -1.0, 0.99)
    g.set_periodic_pressure(0, 1.01, 0.99)
    g.set_spheres(np.array([[4.0, 4.0, 4.0]]), np.zeros((1, 3)), np.zeros((1, 3)), np.array([2.0]),
                  np.array([1]))
    g.diff_forward(2, 1)
    g.set_periodic_pressure(0, 1.0, 1.0, enable=False)
    with pytest.raises(L.PSMError):
        g.diff_backward(np.zeros((2, 1, 6)))             # the step map changed since the forward


def test_macroscopic_fields_match_populations():
    n = (10, 8, 12)
    g = L.Lattice(*n, periodic=(1, 1, 1), tau=0.8, body_accel=(0, 0, 1e-4))
    rho0, u0 = random_state(n, 5)
    g.init_equilibrium(rho0, u0)
    g.collide_stream(2)
    rho, u = g.macroscopic()
    r_h, u_h = L.macroscopic(g.get_populations(), body_accel=(0, 0, 1e-4))
    assert np.abs(rho - r_h).max() < 1e-15
    assert np.abs(u - u_h).max() < 1e-15


@pytest.mark.parametrize("precision", ["fp32_dev", "fp16_store"])
def test_host_and_perturbed_records_collide_like_voxelized(precision):
    """With FP32/FP16 storage the collide reads the float view of the
    records.  Records set from the host, and records whose fraction was
    perturbed, must collide exactly as records the fold wrote with the same
    values."""
    n = (20, 18, 22)
    x = np.array([[6.2, 7.1, 8.3], [13.6, 10.4, 14.8]])
    v = np.array([[0.004, 0.0, 0.001], [-0.002, 0.003, 0.0]])
    w = np.array([[0.0, 0.001, 0.0], [0.0, 0.0, -0.001]])
    r = np.array([3.5, 3.0])
    ids = np.array([1, 2])
    rho, u = random_state(n, 12)
    a = L.Lattice(*n, periodic=(1, 1, 1), tau=0.7, slices_per_dim=5, precision=precision)
    b = L.Lattice(*n, periodic=(1, 1, 1), tau=0.7, slices_per_dim=5, precision=precision)
    for lat in (a, b):
        lat.init_equilibrium(rho, u)
    a.set_spheres(x, v, w, r, ids)
    rec = a.cell_records()
    b.set_cell_records(rec["sf"], rec["id"], rec["us"])
    a.collide_stream(1)
    b.collide_stream(1)
    assert np.array_equal(a.get_populations(), b.get_populations())
    # a perturbed fraction (partial cells only, staying inside (0.001, 1])
    partial = (rec["sf"] > 0.1) & (rec["sf"] < 0.9)
    assert partial.sum() > 10
    dsf = np.where(partial, 0.05, 0.0)
    a.perturb_solid_fraction(dsf)
    b.set_cell_records(rec["sf"] + dsf, rec["id"], rec["us"])
    a.collide_stream(1)
    b.collide_stream(1)
    assert np.array_equal(a.get_populations(), b.get_populations())
    a.close()
    b.close()


def test_checkpoint_save_load_continues_identically(tmp_path):
    n = (20, 18, 22)
    x = np.array([[6.2, 7.1, 8.3], [13.6, 10.4, 14.8]])
    v = np.array([[0.004, 0.0, 0.001], [-0.002, 0.003, 0.0]])
    w = np.array([[0.0, 0.001, 0.0], [0.0, 0.0, -0.001]])
    r = np.array([3.5, 3.0])
    ids = np.array([1, 2])
    a = L.Lattice(*n, periodic=(1, 1, 1), tau=0.7, slices_per_dim=5)
    a.init_equilibrium(*random_state(n, 8))
    a.set_spheres(x, v, w, r, ids)
    a.collide_stream(3)
    a.save_checkpoint(tmp_path / "s.bin")
    b = L.Lattice(*n, periodic=(1, 1, 1), tau=0.7, slices_per_dim=5)
    b.load_checkpoint(tmp_path / "s.bin")
    assert np.array_equal(a.get_populations(), b.get_populations())
    ra, rb = a.cell_records(), b.cell_records()
    for k in ("sf", "id", "us"):
        assert np.array_equal(ra[k], rb[k])
    # both continue with the same particles: identical populations and forces
    for lat in (a, b):
        lat.set_spheres(x, v, w, r, ids)
        lat.collide_stream(2)
    assert np.array_equal(a.get_populations(), b.get_populations())
    Fa, Ta = a.sum_force_torque(x)
    Fb, Tb = b.sum_force_torque(x)
    assert np.array_equal(Fa, Fb) and np.array_equal(Ta, Tb)
    # the reference's writeVTK fields of a lattice
    from test_io import read_vti
    a.write_vti(tmp_path / "a.vti", spacing=2.0, vel_scale=3.0)
    nn, sp, got = read_vti(tmp_path / "a.vti")
    rho, u = a.macroscopic()
    sf = a.cell_records()["sf"]
    c = (4 * n[1] + 5) * n[2] + 6
    assert nn == n and sp == 2.0
    assert got["density"][6, 5, 4, 0] == np.float32(rho[c])
    assert got["pressure"][6, 5, 4, 0] == np.float32(rho[c] - 1.0)
    assert np.array_equal(got["velocity"][6, 5, 4], (3.0 * u[:, c]).astype(np.float32))
    assert got["solidfraction"][6, 5, 4, 0] == np.float32(sf[c])
