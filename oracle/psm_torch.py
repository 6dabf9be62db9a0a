"""Differentiable FP64 restatement of the frozen-geometry PSM window (torch, CPU).

TEST INFRASTRUCTURE ONLY: the reference gradient the hand-written HIP
adjoint is checked against.  Only tests/ import it.

It restates, vectorised over cells, the same cell update as
oracle/psm_oracle.cpp (src/ibCompositeDynamics3D.hh:130-203 with the
textbook D3Q19 BGK / Guo forcing and push streaming with half-way
bounce-back) and the force/torque sum of SumForceTorque3D
(src/ibProcessors3D.hh:202-275) with torque arms about the owner's centre
under the minimum image, and lets torch.autograd differentiate it.  The
geometry is frozen: which cells carry a record and which particle owns each
record are inputs; the record values are functions of the differentiable
parameters:
  sf_c = sf0_c + dsf_c               (sub-cell count mode: sf0 fixed), or
  sf_c = overlap(|d_c|, r_o) + dsf_c (analytic overlap mode),
  us_c = v_o + w_o x d_c,  d_c = minimum image of (cell - x_o).
Its forward is pinned against the C oracle in tests/test_psm_torch.py.
"""
from __future__ import annotations

import math

import numpy as np
import torch

Q = 19
C_NP = np.array([
    [0, 0, 0], [-1, 0, 0], [0, -1, 0], [0, 0, -1], [-1, -1, 0], [-1, 1, 0], [-1, 0, -1],
    [-1, 0, 1], [0, -1, -1], [0, -1, 1], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 0],
    [1, -1, 0], [1, 0, 1], [1, 0, -1], [0, 1, 1], [0, 1, -1]], dtype=np.int64)
W_NP = np.array([1 / 3] + [1 / 18] * 3 + [1 / 36] * 6 + [1 / 18] * 3 + [1 / 36] * 6)
OPP = np.array([0] + list(range(10, 19)) + list(range(1, 10)), dtype=np.int64)
SOLFRAC_MIN = 0.001   # src/ibDef.h:26
SOLFRAC_MAX = 0.999   # src/ibDef.h:27
RC = (3.0 / (4.0 * math.pi)) ** (1.0 / 3.0)
VC = 4.0 / 3.0 * math.pi * RC ** 3

C = torch.tensor(C_NP, dtype=torch.float64)
W = torch.tensor(W_NP, dtype=torch.float64)


def stream_index(n, periodic):
    """Flat gather index: f_new.view(-1) = post.view(-1)[idx] reproduces push
    streaming f_new[x + c_i, i] = post[x, i] with half-way bounce-back
    f_new[x, opp(i)] = post[x, i] where x + c_i leaves a wall axis."""
    nx, ny, nz = n
    X, Y, Z = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    X, Y, Z = X.ravel(), Y.ravel(), Z.ravel()
    nc = nx * ny * nz
    idx = np.empty((nc, Q), dtype=np.int64)
    for i in range(Q):
        src = [X - C_NP[i, 0], Y - C_NP[i, 1], Z - C_NP[i, 2]]
        wall = np.zeros(nc, dtype=bool)
        for d in range(3):
            out = (src[d] < 0) | (src[d] >= n[d])
            if periodic[d]:
                src[d] = np.mod(src[d], n[d])
            else:
                wall |= out
        cell = (src[0] * ny + src[1]) * nz + src[2]
        own = (X * ny + Y) * nz + Z
        idx[:, i] = np.where(wall, own * Q + OPP[i], cell * Q + i)
    return torch.from_numpy(idx)


def equilibria(rho_bar, inv_rho, j):
    cj = j @ C.T
    jsq = (j * j).sum(1, keepdim=True)
    return W * ((rho_bar[:, None] + 3.0 * cj) + 1.5 * inv_rho[:, None] * (3.0 * cj * cj - jsq))


def bgk(f, omega, acc):
    rho_bar = f.sum(1)
    j = f @ C
    rho = rho_bar + 1.0
    inv_rho = 1.0 / rho
    if acc is None:
        return f * (1.0 - omega) + omega * equilibria(rho_bar, inv_rho, j)
    u = j * inv_rho[:, None] + 0.5 * acc
    jeq = rho[:, None] * u
    out = f * (1.0 - omega) + omega * equilibria(rho_bar, inv_rho, jeq)
    pref = (1.0 - 0.5 * omega) * rho
    ca = (acc @ C.T)[None, :]
    cu = u @ C.T
    ua = (u * acc).sum(1, keepdim=True)
    g = 3.0 * (ca - ua) + 9.0 * (cu * ca)
    return out + pref[:, None] * W * g


def overlap(d, r):
    """Analytic overlap fraction (same formula as orc_overlap_fraction)."""
    outside = d >= r + RC
    inside = d <= torch.abs(r - RC)
    ds = torch.where(outside | inside, torch.ones_like(d), d)
    t = (r + RC) - ds
    rm = r - RC
    poly = (ds * ds + 2.0 * ds * (r + RC)) - 3.0 * rm * rm
    lens = (math.pi * t * t * poly) / (12.0 * ds) / VC
    full = torch.where(r >= RC, torch.ones_like(r), (4.0 / 3.0 * math.pi) * r ** 3 / VC)
    return torch.where(outside, torch.zeros_like(d), torch.where(inside, full, lens))


class FrozenWindow:
    """The window's fixed data: lattice, records' owners, streaming index."""

    def __init__(self, n, periodic, live, owner, acc=None, sf0=None):
        self.n = tuple(int(v) for v in n)
        self.periodic = tuple(int(bool(p)) for p in periodic)
        self.nc = self.n[0] * self.n[1] * self.n[2]
        self.idx = stream_index(self.n, self.periodic)
        self.live = torch.as_tensor(np.asarray(live, dtype=bool))
        self.owner = torch.as_tensor(np.asarray(owner, dtype=np.int64))
        self.acc = None if acc is None or not np.any(acc) else torch.tensor(acc, dtype=torch.float64)
        self.sf0 = None if sf0 is None else torch.as_tensor(np.asarray(sf0, dtype=np.float64))
        X, Y, Z = np.meshgrid(*[np.arange(v) for v in self.n], indexing="ij")
        self.cell = torch.tensor(np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1), dtype=torch.float64)
        self.lc = torch.nonzero(self.live).ravel()          # live cells
        self.lo = self.owner[self.lc]                        # their owners

    def arms(self, x):
        d = self.cell[self.lc] - x[self.lo]
        nn = torch.tensor(self.n, dtype=torch.float64)
        half = torch.tensor([v // 2 for v in self.n], dtype=torch.float64)
        d = torch.where(d > half, d - nn, torch.where(d < -half, d + nn, d))
        return d

    def run(self, f0, x, v, w, r, tau, dsf, nsteps, mode):
        """Forces and torques per particle per step, (nsteps, n, 6)."""
        npart = x.shape[0]
        omega = 1.0 / tau
        ooo = 1.0 / omega - 0.5
        d = self.arms(x)
        if mode == "overlap":
            dist = torch.sqrt((d * d).sum(1))
            sf_l = overlap(dist, r[self.lo])
        else:
            sf_l = self.sf0[self.lc]
        sf_l = sf_l + dsf[self.lc]
        us_l = v[self.lo] + torch.cross(w[self.lo], d, dim=1)
        full = sf_l > SOLFRAC_MAX
        B = torch.where(full, torch.ones_like(sf_l), sf_l * ooo / ((1.0 - sf_l) + ooo))
        f = f0
        out = []
        for _ in range(nsteps):
            fb = bgk(f, omega, self.acc)
            fP = f[self.lc]
            fbl = torch.where(full[:, None], fP, fb[self.lc])
            rho_b = fbl.sum(1)
            jb = fbl @ C
            inv = 1.0 / (rho_b + 1.0)
            E = equilibria(rho_b, inv, jb)
            S = equilibria(rho_b, inv, us_l * (1.0 + rho_b)[:, None])
            D = (fP[:, OPP] - E[:, OPP]) - (fP - S)
            D = torch.cat([(E[:, :1] - S[:, :1]), D[:, 1:]], 1)
            post_l = (fP + (1.0 - B)[:, None] * (fbl - fP)) + B[:, None] * D
            F_l = -((B[:, None] * D) @ C)
            post = fb.index_put((self.lc,), post_l)
            f = post.reshape(-1)[self.idx.reshape(-1)].reshape(self.nc, Q)
            Fp = torch.zeros(npart, 3, dtype=torch.float64).index_add(0, self.lo, F_l)
            Tp = torch.zeros(npart, 3, dtype=torch.float64).index_add(0, self.lo, torch.cross(d, F_l, dim=1))
            out.append(torch.cat([Fp, Tp], 1))
        return torch.stack(out), f
