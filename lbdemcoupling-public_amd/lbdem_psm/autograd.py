"""torch.autograd.Function over the differentiable PSM window.

``psm_window(lattice, f0, x, v, w, r, ids, tau, dsf, nsteps, checkpoint_every)``
runs, on the GPU through the C ABI:

1. ``set_tau(tau)``, ``set_populations(f0)``, ``clear_records()`` and
   ``set_spheres(x, v, w, r, ids)`` (a fresh ownership history, so that the
   function depends on its arguments only),
2. ``perturb_solid_fraction(dsf)`` when ``dsf`` is non-zero,
3. ``diff_forward(nsteps, checkpoint_every)``,

and returns the per-step particle forces and torques, shape (nsteps, n, 6).
The backward pass is the library's hand-written adjoint
(``lbdem_psm_diff_backward``): gradients with respect to x (through the
analytic overlap when the lattice uses ``voxel_mode="overlap"``, and the
lever and torque arms), v, w, tau and the per-cell solid fraction ``dsf``.
Torch is used for the autograd plumbing only; tensors live on the host.
"""
from __future__ import annotations

import numpy as np
import torch

from . import Lattice


def _np(t):
    return t.detach().cpu().numpy().astype(np.float64, copy=False)


class PSMWindow(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, v, w, tau, dsf, lattice: Lattice, f0, r, ids, nsteps, checkpoint_every,
                auto_images):
        lattice.set_tau(float(tau))
        lattice.set_populations(f0)
        lattice.clear_records()
        lattice.set_spheres(_np(x), _np(v), _np(w), r, ids, auto_images=auto_images)
        d = _np(dsf)
        if np.any(d != 0.0):
            lattice.perturb_solid_fraction(d)
        ft = lattice.diff_forward(int(nsteps), int(checkpoint_every))
        ctx.lattice = lattice
        ctx.token = lattice.window_token()
        ctx.dtype = x.dtype
        return torch.from_numpy(ft).to(x.dtype)

    @staticmethod
    def backward(ctx, g_ft):
        # raises if the lattice ran another window after this forward pass
        # (its checkpoints, tau and records would be that window's)
        g = ctx.lattice.diff_backward(g_ft=_np(g_ft.contiguous()), token=ctx.token)
        dt = ctx.dtype
        return (torch.from_numpy(g["x"]).to(dt), torch.from_numpy(g["v"]).to(dt),
                torch.from_numpy(g["w"]).to(dt), torch.tensor(g["tau"], dtype=dt),
                torch.from_numpy(g["sf"]).to(dt), None, None, None, None, None, None, None)


def psm_window(lattice: Lattice, f0, x, v, w, r, ids, tau, dsf=None, nsteps=1,
               checkpoint_every=1, auto_images=True):
    """Per-step (F, T) of every particle over a frozen-geometry window."""
    if dsf is None:
        dsf = torch.zeros(lattice.nc, dtype=torch.float64)
    if not torch.is_tensor(tau):
        tau = torch.tensor(float(tau), dtype=torch.float64)
    return PSMWindow.apply(x, v, w, tau, dsf, lattice, f0, r, ids, nsteps, checkpoint_every,
                           auto_images)
