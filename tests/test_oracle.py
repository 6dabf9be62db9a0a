"""Pins of the CPU oracle (oracle/psm_oracle.cpp) against closed forms,
invariants, brute force and hand-derived cases of the reference rules.

The reference has no tests, fixtures or golden vectors of its own; every
check here is chosen so that a plausible mistake (a wrong sign or index, a
dropped term, a swapped operand, the wrong pass order) fails one of them.
"""
import itertools

import numpy as np
import pytest

from oracle import oracle as O

C = np.array([
    [0, 0, 0], [-1, 0, 0], [0, -1, 0], [0, 0, -1], [-1, -1, 0], [-1, 1, 0], [-1, 0, -1],
    [-1, 0, 1], [0, -1, -1], [0, -1, 1], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 0],
    [1, -1, 0], [1, 0, 1], [1, 0, -1], [0, 1, 1], [0, 1, -1]], dtype=np.float64)
W = np.array([1 / 3] + [1 / 18] * 3 + [1 / 36] * 6 + [1 / 18] * 3 + [1 / 36] * 6)


def textbook_eq(rho, u):
    """f_i^eq - t_i from the textbook form t_i rho (1 + 3 c.u + 4.5 (c.u)^2 - 1.5 u^2)."""
    cu = C @ u
    return W * rho * (1 + 3 * cu + 4.5 * cu ** 2 - 1.5 * u @ u) - W


def test_lattice_constants():
    assert np.allclose(C[1:10], -C[10:19])
    assert abs(W.sum() - 1) < 1e-15
    assert np.allclose((W[:, None, None] * C[:, :, None] * C[:, None, :]).sum(0), np.eye(3) / 3)


@pytest.mark.parametrize("seed", range(5))
def test_equilibrium_moments_and_textbook_form(seed):
    rng = np.random.default_rng(seed)
    rho_bar = rng.normal(0, 0.05)
    j = rng.normal(0, 0.03, 3)
    eq = O.equilibrium(rho_bar, j)
    rho = 1 + rho_bar
    assert abs(eq.sum() - rho_bar) < 1e-15
    assert np.allclose(C.T @ eq, j, atol=1e-16, rtol=0)
    Pi = (eq[:, None, None] * C[:, :, None] * C[:, None, :]).sum(0)
    assert np.allclose(Pi, rho_bar / 3 * np.eye(3) + np.outer(j, j) / rho, atol=1e-16, rtol=0)
    assert np.allclose(eq, textbook_eq(rho, j / rho), atol=1e-16, rtol=0)


def random_cell(rng, scale=0.01):
    rho_bar = rng.normal(0, 0.02)
    j = rng.normal(0, 0.02, 3)
    return O.equilibrium(rho_bar, j) + rng.normal(0, scale, 19) * W


@pytest.mark.parametrize("seed", range(4))
def test_fluid_cell_is_textbook_bgk(seed):
    rng = np.random.default_rng(seed)
    f = random_cell(rng)
    tau = 0.6 + seed * 0.2
    out, F = O.collide_cell(f, tau, 0.0, np.zeros(3))
    rho = 1 + f.sum()
    u = (C.T @ f) / rho
    ref = f - (f - textbook_eq(rho, u)) / tau
    assert np.allclose(out, ref, atol=1e-15, rtol=0)
    assert np.all(F == 0)
    assert abs(out.sum() - f.sum()) < 1e-15
    assert np.allclose(C.T @ out, C.T @ f, atol=1e-16, rtol=0)


@pytest.mark.parametrize("sf", [0.0016, 0.13, 0.5, 0.87, 0.9985, 1.0])
@pytest.mark.parametrize("seed", range(3))
def test_momentum_exchange_identity(sf, seed):
    """F = j_in - j_out: the force on the particle is the momentum the cell's
    fluid lost (BGK keeps j), for every solid fraction."""
    rng = np.random.default_rng(100 + seed)
    f = random_cell(rng)
    us = rng.normal(0, 0.02, 3)
    out, F = O.collide_cell(f, 0.8, sf, us)
    assert np.allclose(F, C.T @ f - C.T @ out, atol=2e-16, rtol=0)


@pytest.mark.parametrize("seed", range(3))
def test_full_solid_cell_resets_to_particle_equilibrium(seed):
    """sf > 0.999, f at equilibrium with velocity u: the moving populations
    leave with the equilibrium of the particle velocity and F = j - rho u_s.
    The rest population gets fP_0 + (E_0 - Es_0) = 2 E_0 - Es_0: the
    reference adds fEq[0] - fEqSolid[0] (src/ibCompositeDynamics3D.hh:161-162),
    the opposite sign of the solid term of the moving pairs."""
    rng = np.random.default_rng(200 + seed)
    rho_bar = rng.normal(0, 0.02)
    j = rng.normal(0, 0.02, 3)
    us = rng.normal(0, 0.02, 3)
    f = O.equilibrium(rho_bar, j)
    out, F = O.collide_cell(f, 0.7, 1.0, us)
    rho = 1 + rho_bar
    Es = O.equilibrium(rho_bar, rho * us)
    assert np.allclose(out[1:], Es[1:], atol=1e-16, rtol=0)
    assert abs(out[0] - (2 * f[0] - Es[0])) < 1e-17
    assert np.allclose(F, j - rho * us, atol=1e-16, rtol=0)


@pytest.mark.parametrize("sf", [0.2, 0.7, 1.0])
def test_comoving_solid_exerts_no_force(sf):
    rng = np.random.default_rng(3)
    rho_bar = 0.01
    u = rng.normal(0, 0.02, 3)
    f = O.equilibrium(rho_bar, (1 + rho_bar) * u)
    out, F = O.collide_cell(f, 0.9, sf, u)
    assert np.allclose(out, f, atol=1e-16, rtol=0)
    assert np.allclose(F, 0, atol=1e-16)


def test_weighting_interpolates_fluid_and_solid():
    """Linearity in B: out(sf) = out(0) + B(sf) * (out_solid_rule(B=1 with BGK) - out(0))
    with B = sf (tau-1/2)/((1-sf)+(tau-1/2)) (src/ibCompositeDynamics3D.hh:177-179)."""
    rng = np.random.default_rng(9)
    f = random_cell(rng)
    us = rng.normal(0, 0.02, 3)
    tau = 0.8
    out0, _ = O.collide_cell(f, tau, 0.0, us)
    for sf in (0.1, 0.4, 0.8):
        B = sf * (tau - 0.5) / ((1 - sf) + (tau - 0.5))
        outs, Fs = O.collide_cell(f, tau, sf, us)
        # the B-part is B*Omega_s with Omega_s built from the post-BGK moments
        omega_s = (outs - out0 - B * (f - out0)) / B   # out = fP + (1-B)(fB - fP) + B*Om
        sf2 = 0.6
        B2 = sf2 * (tau - 0.5) / ((1 - sf2) + (tau - 0.5))
        outs2, _ = O.collide_cell(f, tau, sf2, us)
        pred = out0 + B2 * (f - out0) + B2 * omega_s
        assert np.allclose(outs2, pred, atol=1e-15, rtol=0)


def test_pse_mass_defect_of_rest_population():
    """The reference's rest-population rule changes a solid cell's mass by
    2 B (E_0 - Es_0) (pairs i / i+9 cancel in pairs, the rest term does not)."""
    rng = np.random.default_rng(11)
    f = random_cell(rng)
    us = rng.normal(0, 0.03, 3)
    tau = 0.7
    sf = 0.45
    out, _ = O.collide_cell(f, tau, sf, us)
    fB, _ = O.collide_cell(f, tau, 0.0, us)
    rho_bar = fB.sum()
    j = C.T @ fB
    E0 = O.equilibrium(rho_bar, j)[0]
    Es0 = O.equilibrium(rho_bar, (1 + rho_bar) * us)[0]
    B = sf * (tau - 0.5) / ((1 - sf) + (tau - 0.5))
    assert abs((out.sum() - f.sum()) - 2 * B * (E0 - Es0)) < 1e-16


# ---------------------------------------------------------------- voxelization

def brute_count(dx, dy, dz, r, N):
    w = 1.0 / N
    delta = -0.5 + (np.arange(N, dtype=np.float64) + 0.5) * w
    ax, ay, az = dx + delta, dy + delta, dz + delta
    xs, ys, zs = ax * ax, ay * ay, az * az
    s = (xs[:, None, None] + ys[None, :, None]) + zs[None, None, :]
    return int((s < r * r).sum())


@pytest.mark.parametrize("N", [5, 8])
def test_subcell_count_matches_brute_force(N):
    rng = np.random.default_rng(N)
    for _ in range(3000):
        r = rng.uniform(2.5, 9.0)
        d = rng.normal(0, 1, 3)
        d *= (r + rng.uniform(-1.2, 1.2)) / np.linalg.norm(d)
        sf, n, br = O.calc_solid_fraction(*d, r, N)
        bc = brute_count(*d, r, N)
        if br == 2:
            assert n == bc
            assert sf == (1.0 / N ** 3) * float(n)
        elif br == 0:
            assert bc == 0 and sf == 0.0 and n == 0
        else:
            assert bc == N ** 3 and sf == 1.0 and n == N ** 3


def test_subcell_shortcut_radius():
    h = np.sqrt(3.1) / 2
    r = 5.0
    assert O.calc_solid_fraction(r + h + 1e-9, 0, 0, r, 5)[2] == 0
    assert O.calc_solid_fraction(r + h - 1e-9, 0, 0, r, 5)[2] == 2
    assert O.calc_solid_fraction(r - h - 1e-9, 0, 0, r, 5)[2] == 1
    assert O.calc_solid_fraction(r - h + 1e-9, 0, 0, r, 5)[2] == 2


@pytest.mark.parametrize("N", [5, 8])
def test_subcell_counts_cube_symmetric(N):
    r = 4.3
    for d in itertools.product(range(-6, 7), repeat=3):
        base = O.calc_solid_fraction(*map(float, d), r, N)[1]
        for perm in itertools.permutations(range(3)):
            for sgn in itertools.product((1, -1), repeat=3):
                e = [sgn[k] * d[perm[k]] for k in range(3)]
                assert O.calc_solid_fraction(*map(float, e), r, N)[1] == base
        if d[0] > 0:
            break


@pytest.mark.parametrize("N", [5, 8])
def test_sphere_volume_converges(N):
    r = 6.0
    ol = O.OracleLattice(24, 24, 24, tau=0.8, slices_per_dim=N)
    x = np.array([[11.3, 12.1, 11.7]])
    ol.set_spheres(x, np.zeros((1, 3)), np.zeros((1, 3)), [r], [1])
    vol = ol.cell_records()["sf"].sum()
    assert abs(vol - 4 / 3 * np.pi * r ** 3) / (4 / 3 * np.pi * r ** 3) < (0.01 if N == 5 else 0.003)


def test_bounding_box_truncates_toward_zero():
    """Box3D(x-r-2, x+r+2) converts with truncation (src/ibProcessors3D.h:43-45):
    a far-reaching sphere near the lower face still only touches cells >= 0,
    and a cell at index trunc(x+r+2)+1 is never visited (it keeps its record)."""
    ol = O.OracleLattice(16, 16, 16, periodic=(0, 0, 0), tau=0.8, slices_per_dim=5)
    sf = np.zeros(16 ** 3)
    ids = np.zeros(16 ** 3, dtype=np.int64)
    # plant a stale record just outside the box of the sphere below
    x = np.array([[5.25, 8.0, 8.0]])
    r = 2.0
    hi = int(5.25 + r + 2)          # = 9
    stale = ((hi + 1) * 16 + 8) * 16 + 8
    sf[stale] = 0.5
    ids[stale] = 1
    ol.set_cell_records(sf, ids, np.zeros((3, 16 ** 3)))
    ol.set_spheres(x, np.zeros((1, 3)), np.zeros((1, 3)), [r], [1])
    rec = ol.cell_records()
    assert rec["sf"][stale] == 0.5 and rec["id"][stale] == 1    # not visited
    inside = ((hi) * 16 + 8) * 16 + 8
    assert rec["sf"][inside] == 0.0                              # visited and zeroed


# ------------------------------------------------------------ ownership rule

def _two_sphere_setup(order, prior=None):
    """Two spheres whose shells overlap around cell (8,8,8)."""
    n = 16
    ol = O.OracleLattice(n, n, n, periodic=(0, 0, 0), tau=0.8, slices_per_dim=5)
    xa = np.array([8.0 - 2.9, 8.0, 8.0])
    xb = np.array([8.0 + 3.15, 8.0, 8.0])
    X = np.array([xa, xb])[list(order)]
    ids = np.array([1, 2])[list(order)]
    r = np.array([3.0, 3.0])
    if prior is not None:
        sf = np.zeros(n ** 3)
        pid = np.zeros(n ** 3, dtype=np.int64)
        c = (8 * n + 8) * n + 8
        sf[c], pid[c] = prior
        ol.set_cell_records(sf, pid, np.zeros((3, n ** 3)))
    ol.set_spheres(X, np.zeros((2, 3)), np.zeros((2, 3)), r, ids, auto_images=False)
    c = (8 * n + 8) * n + 8
    rec = ol.cell_records()
    sfa = O.calc_solid_fraction(8 - xa[0], 0, 0, 3.0, 5)[0]
    sfb = O.calc_solid_fraction(8 - xb[0], 0, 0, 3.0, 5)[0]
    return rec["sf"][c], rec["id"][c], sfa, sfb


def test_ownership_larger_fraction_wins_both_orders():
    sf1, id1, sfa, sfb = _two_sphere_setup((0, 1))
    sf2, id2, _, _ = _two_sphere_setup((1, 0))
    assert 0 < sfb < sfa
    assert (id1, sf1) == (1, sfa)        # 1 takes (case 1), 2 smaller (case 3 no)
    assert (id2, sf2) == (1, sfa)        # 2 takes, 1 larger (case 3 yes)


def test_ownership_history_owner_keeps_cell():
    """Previous owner 2 with sf 0.9: in order (1, 2) sphere 1 (sfa < 0.9) does
    not take it, sphere 2 re-takes it by id (case 3, id_old == id) with its
    smaller fraction; in order (2, 1) sphere 2 re-takes it, then sphere 1 with
    the larger fraction takes it."""
    sf1, id1, sfa, sfb = _two_sphere_setup((0, 1), prior=(0.9, 2))
    assert (id1, sf1) == (2, sfb)
    sf2, id2, _, _ = _two_sphere_setup((1, 0), prior=(0.9, 2))
    assert (id2, sf2) == (1, sfa)


def test_ownership_leaving_particle_clears_cell():
    """Case 2: a cell owned by id 7 with no particle fraction left is cleared
    by particle 7's own pass (id match) but kept by another particle's pass."""
    n = 16
    c = (2 * n + 2) * n + 2
    for pid_in_cell, expect_cleared in ((7, True), (9, False)):
        ol = O.OracleLattice(n, n, n, periodic=(0, 0, 0), tau=0.8, slices_per_dim=5)
        sf = np.zeros(n ** 3)
        pid = np.zeros(n ** 3, dtype=np.int64)
        sf[c], pid[c] = 0.4, pid_in_cell
        ol.set_cell_records(sf, pid, np.zeros((3, n ** 3)))
        # sphere 7 at distance > r + h from cell (2,2,2) but with the cell in its box
        ol.set_spheres(np.array([[6.0, 6.0, 6.0]]), np.zeros((1, 3)), np.zeros((1, 3)), [2.5], [7],
                       auto_images=False)
        rec = ol.cell_records()
        assert (rec["sf"][c] == 0.0) == expect_cleared


def test_set_values_rigid_body_velocity():
    """u_s = v + omega x (x_cell - x_com) (src/ibProcessors3D.hh:162-174)."""
    n = 16
    ol = O.OracleLattice(n, n, n, tau=0.8, slices_per_dim=5)
    x = np.array([[7.5, 8.25, 7.75]])
    v = np.array([[0.01, -0.02, 0.005]])
    w = np.array([[0.001, 0.002, -0.003]])
    ol.set_spheres(x, v, w, [3.0], [1])
    rec = ol.cell_records()
    idx = np.nonzero(rec["sf"] > 0)[0]
    X = np.stack(np.unravel_index(idx, (n, n, n)), axis=1).astype(np.float64)
    expect = v + np.cross(w, X - x)
    assert np.allclose(rec["us"][:, idx].T, expect, atol=1e-18, rtol=0)


def test_exclude_type_skips_particle():
    n = 16
    ol = O.OracleLattice(n, n, n, tau=0.8, slices_per_dim=5)
    x = np.array([[4.0, 4.0, 4.0], [11.0, 11.0, 11.0]])
    ol.set_spheres(x, np.zeros((2, 3)), np.zeros((2, 3)), [2.0, 2.0], [1, 2], types=[1, 2],
                   exclude=[2])
    ids = ol.cell_records()["id"]
    assert set(np.unique(ids)) == {0, 1}


def test_periodic_image_equals_translated_sphere():
    """A sphere straddling the periodic faces voxelizes like the same sphere
    moved inside by a whole number of cells, cyclically shifted."""
    n = 16
    x_in = np.array([[8.25, 7.5, 8.75]])
    x_wrap = x_in + np.array([[7.0, -8.0, 8.0]])
    x_wrap = np.where(x_wrap >= n - 0.5, x_wrap - n, x_wrap)
    x_wrap = np.where(x_wrap < -0.5, x_wrap + n, x_wrap)
    out = []
    for xx in (x_in, x_wrap):
        ol = O.OracleLattice(n, n, n, tau=0.8, slices_per_dim=8)
        ol.set_spheres(xx, np.zeros((1, 3)), np.zeros((1, 3)), [3.5], [1])
        out.append(ol.cell_records()["sf"].reshape(n, n, n))
    shifted = np.roll(out[0], shift=(7, -8, 8), axis=(0, 1, 2))
    assert np.array_equal(shifted, out[1])


# ------------------------------------------------------------ force / torque

def brute_force_torque(rec, n, x, ids):
    F = np.zeros((len(x), 3))
    T = np.zeros((len(x), 3))
    idmap = {int(i): k for k, i in reversed(list(enumerate(ids)))}
    sfid = rec["id"]
    for c in range(n[0] * n[1] * n[2]):
        pid = int(sfid[c])
        if pid < 1:
            continue
        k = idmap[pid]
        X = np.unravel_index(c, n)
        d = [float(X[a]) - x[k][a] for a in range(3)]
        for a in range(3):
            if d[a] > float(n[a] // 2):
                d[a] -= n[a]
            elif d[a] < float(-(n[a] // 2)):
                d[a] += n[a]
        f = rec["F"][:, c]
        F[k] += f
        T[k] += np.array([d[1] * f[2] - d[2] * f[1], -d[0] * f[2] + d[2] * f[0],
                          d[0] * f[1] - d[1] * f[0]])
    return F, T


def test_force_torque_sum_matches_brute_force_with_minimum_image():
    n = (20, 18, 16)
    ol = O.OracleLattice(*n, tau=0.7, slices_per_dim=5)
    rng = np.random.default_rng(5)
    ol.init_equilibrium(u=np.tile(np.array([0.01, -0.005, 0.002])[:, None], (1, np.prod(n))))
    x = np.array([[0.2, 9.0, 8.0], [10.0, 17.1, 3.0], [12.0, 6.0, 15.2]])   # all straddle a face
    v = rng.normal(0, 0.01, (3, 3))
    w = rng.normal(0, 1e-3, (3, 3))
    ol.set_spheres(x, v, w, [3.0, 2.5, 3.3], [1, 2, 3])
    ol.collide_stream(1)
    F, T = ol.sum_force_torque(x)
    Fb, Tb = brute_force_torque(ol.cell_records(), n, x, [1, 2, 3])
    assert np.allclose(F, Fb, rtol=1e-13, atol=1e-16)
    assert np.allclose(T, Tb, rtol=1e-13, atol=1e-15)
    assert np.abs(F).max() > 1e-4


def test_slab_threads_give_the_sequential_records_and_sums():
    """The x-slab form of the particle loop (one slab per thread, the
    reference's per-rank loop) gives the sequential records bit for bit,
    including ownership history and overlapping boxes; the slab-wise force
    sums agree to rounding."""
    n = (23, 16, 18)
    out = []
    for threads in (1, 4):
        ol = O.OracleLattice(*n, tau=0.7, slices_per_dim=5)
        ol.set_threads(threads)
        rng = np.random.default_rng(8)
        x = np.array([[3.0, 4.0, 5.0], [8.6, 5.2, 6.1], [15.0, 12.0, 0.3], [21.8, 8.0, 9.0]])
        r = [3.1, 2.8, 3.6, 2.5]
        v = rng.normal(0, 0.01, (4, 3))
        w = rng.normal(0, 1e-3, (4, 3))
        recs = []
        for s in range(3):
            ol.set_spheres(x + 0.37 * s, v, w, r, [1, 2, 3, 4])
            ol.collide_stream(1, 1)
            recs.append(ol.cell_records())
        F, T = ol.sum_force_torque(x + 0.74)
        out.append((recs, F, T))
    for a, b in zip(out[0][0], out[1][0]):
        for k in ("sf", "id", "us", "F"):
            assert np.array_equal(a[k], b[k]), k
    assert np.allclose(out[0][1], out[1][1], rtol=1e-12, atol=1e-16)
    assert np.allclose(out[0][2], out[1][2], rtol=1e-12, atol=1e-15)
    assert np.abs(out[0][1]).max() > 1e-5


# ------------------------------------------------------------ streaming, walls, forcing

def test_periodic_fluid_conserves_mass_and_momentum():
    n = (10, 9, 8)
    ol = O.OracleLattice(*n, tau=0.62, slices_per_dim=5)
    rng = np.random.default_rng(1)
    nc = int(np.prod(n))
    rho = 1 + rng.normal(0, 0.01, nc)
    u = rng.normal(0, 0.02, (3, nc))
    ol.init_equilibrium(rho, u)
    m0, j0 = ol.conservation()
    ol.collide_stream(25)
    m1, j1 = ol.conservation()
    assert abs(m1 - m0) < 1e-12
    assert np.allclose(j1, j0, atol=1e-12)


def test_streaming_is_a_permutation():
    """With omega = 1 on a state at equilibrium... use tau = 1 and a state that is
    an equilibrium per cell: collision is the identity, so one step is a pure
    shift of each population by c_i."""
    n = (6, 5, 7)
    ol = O.OracleLattice(*n, tau=1.0, slices_per_dim=5)
    rng = np.random.default_rng(2)
    nc = int(np.prod(n))
    f = np.zeros((19, nc))
    for c in range(nc):
        f[:, c] = O.equilibrium(rng.normal(0, 0.01), rng.normal(0, 0.01, 3))
    ol.set_populations(f)
    ol.collide_stream(1)
    g = ol.get_populations()
    fr = f.reshape(19, *n)
    for i in range(19):
        expect = np.roll(fr[i], shift=tuple(C[i].astype(int)), axis=(0, 1, 2))
        assert np.allclose(g.reshape(19, *n)[i], expect, atol=1e-17, rtol=0)


def test_bounce_back_walls_conserve_mass():
    n = (8, 6, 6)
    ol = O.OracleLattice(*n, periodic=(0, 0, 1), tau=0.7, slices_per_dim=5)
    rng = np.random.default_rng(4)
    nc = int(np.prod(n))
    ol.init_equilibrium(1 + rng.normal(0, 0.01, nc), rng.normal(0, 0.02, (3, nc)))
    m0, _ = ol.conservation()
    ol.collide_stream(40)
    m1, _ = ol.conservation()
    assert abs(m1 - m0) < 1e-12


def test_poiseuille_closed_form():
    """Walls in x (half-way bounce-back), Guo body force along z: the steady
    profile is u_z(x) = a/(2 nu) (x + 1/2)(nx - 1/2 - x), nu = (tau - 1/2)/3,
    exactly for BGK at tau = 1/2 + sqrt(3/16)."""
    nx = 9
    tau = 0.5 + np.sqrt(3.0 / 16.0)
    a = 1e-5
    ol = O.OracleLattice(nx, 2, 2, periodic=(0, 1, 1), tau=tau, slices_per_dim=5,
                         body_accel=(0.0, 0.0, a))
    ol.collide_stream(6000)
    f = ol.get_populations()
    rho = 1 + f.sum(0)
    uz = ((C.T @ f)[2] / rho + 0.5 * a).reshape(nx, 2, 2)[:, 0, 0]
    nu = (tau - 0.5) / 3
    xs = np.arange(nx, dtype=float)
    exact = a / (2 * nu) * (xs + 0.5) * (nx - 0.5 - xs)
    assert np.allclose(uz, exact, rtol=1e-6, atol=0)


def test_galilean_comoving_spheres_feel_no_force():
    n = (16, 16, 16)
    u = np.array([0.02, -0.01, 0.015])
    ol = O.OracleLattice(*n, tau=0.8, slices_per_dim=5)
    nc = int(np.prod(n))
    ol.init_equilibrium(u=np.tile(u[:, None], (1, nc)))
    x = np.array([[4.3, 5.1, 8.0], [11.7, 10.2, 3.3]])
    V = np.tile(u, (2, 1))
    ol.set_spheres(x, V, np.zeros((2, 3)), [2.7, 3.1], [1, 2])
    f0 = ol.get_populations()
    ol.collide_stream(3)
    F, T = ol.sum_force_torque(x)
    assert np.abs(F).max() < 1e-14 and np.abs(T).max() < 1e-13
    assert np.allclose(ol.get_populations(), f0, atol=1e-15, rtol=0)


def test_define_velocity_sets_particle_equilibrium():
    """initVel: cells with sf > 0.999 get the equilibrium at rhoBar = 0 with
    j = u_s (src/ibCompositeDynamics3D.hh:118-127)."""
    n = (16, 16, 16)
    ol = O.OracleLattice(*n, tau=0.8, slices_per_dim=5)
    x = np.array([[8.0, 8.0, 8.0]])
    v = np.array([[0.01, 0.02, -0.01]])
    ol.set_spheres(x, v, np.zeros((1, 3)), [3.5], [1], init_vel=True)
    f = ol.get_populations()
    c = (8 * 16 + 8) * 16 + 8
    assert np.allclose(f[:, c], O.equilibrium(0.0, v[0]), atol=1e-18, rtol=0)
    far = 0
    assert np.allclose(f[:, far], 0.0, atol=1e-18)


# ---------------------------------------------------------------- boundaries

def _macro(f):
    C = np.array([[0, 0, 0], [-1, 0, 0], [0, -1, 0], [0, 0, -1], [-1, -1, 0], [-1, 1, 0], [-1, 0, -1],
                  [-1, 0, 1], [0, -1, -1], [0, -1, 1], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 0],
                  [1, -1, 0], [1, 0, 1], [1, 0, -1], [0, 1, 1], [0, 1, -1]])
    rho = 1.0 + f.sum(0)
    return rho, (C.T @ f) / rho


def test_poiseuille_initial_field_closed_form():
    """PoiseuilleProfileAndPressureGradient: the density ramps from rho_hi at
    index 0 to rho_lo at index n-1 of the profile size, the velocity is
    u_max p(j) p(k) with p(n/2) = 1 and p(0) = 0, along the axis only."""
    n = (10, 9, 12)
    o = O.OracleLattice(*n, periodic=(1, 0, 0), tau=0.8)
    o.init_poiseuille(0, 1.002, 0.998, 0.02, n3=(n[0] - 1, n[1], n[2]))
    rho, u = _macro(o.get_populations())
    rho, u = rho.reshape(n), u.reshape((3,) + n)
    assert rho[0, 3, 3] == pytest.approx(1.002, abs=1e-14)
    assert rho[n[0] - 2, 3, 3] == pytest.approx(0.998, abs=1e-14)      # index nx-1 of the profile size nx-1
    assert np.allclose(np.diff(rho[:, 4, 5]), -0.004 / (n[0] - 2), atol=1e-14)
    # p(j, 9) p(k, 12) at j = 4.5 is not a lattice point: check a direct value
    pj = 1 - ((4 - 4.5) / 4.5) ** 2
    pk = 1 - ((6 - 6.0) / 6.0) ** 2
    assert u[0, 2, 4, 6] == pytest.approx(0.02 * pj * pk, rel=1e-13)
    assert u[0, 2, 0, 6] == pytest.approx(0.0, abs=1e-16)
    assert np.abs(u[1]).max() < 1e-17 and np.abs(u[2]).max() < 1e-17   # round-off of the moment sums


def test_periodic_pressure_drives_poiseuille_flow():
    """Zhang-Kwok periodic pressure across x in a channel with walls in y:
    the density falls linearly along x and the steady centreline velocity
    matches the plane-Poiseuille value u = g/(2 nu) (H/2)^2 (H = ny with
    half-way bounce-back) for the measured mid-channel gradient
    g = c_s^2 drho/dx / rho, within 3 %; the profile is a parabola."""
    n, tau = (12, 14, 3), 0.8
    o = O.OracleLattice(*n, periodic=(1, 0, 1), tau=tau)
    o.init_equilibrium()
    o.set_periodic_pressure(0, 1.0005, 0.9995, range4=(1, n[1] - 2, 0, n[2] - 1))
    o.collide_stream(6000, 1)
    rho, u = _macro(o.get_populations())
    rho, ux = rho.reshape(n), u[0].reshape(n)
    assert np.all(np.diff(rho[:, 7, 1]) < 0)
    grad = (rho[5, 7, 1] - rho[7, 7, 1]) / 2.0
    nu = (tau - 0.5) / 3.0
    y = np.arange(n[1])
    prof = ux[6, :, 1]
    want = grad / 3.0 / rho[6, 7, 1] / (2 * nu) * (y + 0.5) * (n[1] - 0.5 - y)
    assert prof.max() == pytest.approx(want.max(), rel=0.03)
    fit = np.polyfit(y, prof, 2)
    assert np.abs(np.polyval(fit, y) - prof).max() < 0.01 * prof.max()


def test_periodic_pressure_touches_only_the_planes():
    """The rescale acts on the populations that crossed the periodic
    boundary into the inlet (c_x = +1) and outlet (c_x = -1) planes only."""
    n = (8, 6, 5)
    nc = int(np.prod(n))
    res = []
    for pp in (False, True):
        o = O.OracleLattice(*n, periodic=(1, 1, 1), tau=0.7)
        o.init_equilibrium(1 + 0.01 * np.random.default_rng(1).normal(size=nc),
                           0.02 * np.random.default_rng(2).normal(size=(3, nc)))
        if pp:
            o.set_periodic_pressure(0, 1.01, 0.99)
        o.collide_stream(1, 1)
        res.append(o.get_populations().reshape(19, *n))
    d = res[1] != res[0]
    cx = np.array([0, -1, 0, 0, -1, -1, -1, -1, 0, 0, 1, 0, 0, 1, 1, 1, 1, 0, 0])
    assert d.any()
    for i in range(19):
        for x in range(n[0]):
            changed = d[i, x].any()
            if changed:
                assert (x == 0 and cx[i] == 1) or (x == n[0] - 1 and cx[i] == -1)
            if cx[i] == 1 and x == 0 or cx[i] == -1 and x == n[0] - 1:
                # the default box is 1..n-2 on the other axes: its interior changed, its rim did not
                assert d[i, x, 1:-1, 1:-1].all() and not d[i, x, 0, :].any()
