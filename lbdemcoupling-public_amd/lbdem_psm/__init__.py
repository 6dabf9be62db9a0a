"""Python binding of the lbdem_psm C ABI (include/lbdem_psm.h) via ctypes.

The binding mirrors the reference's coupling surface for the PSM hot path:

* ``Lattice``            -- the Palabos lattice whose cells all carry
  ``IBcompositeDynamics(new BGKdynamics(omega))``
  (examples/sedimentingSphere/sedimentingSphere.cpp:39,181).
* ``Lattice.set_spheres`` -- ``setSpheresOnLattice`` + ``AttributeFunctional``
  (src/ibDataExchangeWrappers3D.hh:33-156).
* ``Lattice.collide_stream`` -- ``lattice.collideAndStream()`` with
  ``IBcompositeDynamics::collide`` (src/ibCompositeDynamics3D.hh:130-203).
* ``Lattice.sum_force_torque`` -- ``SumForceTorque3D`` + the ``map(tag)`` fold of
  ``getForcesFromLattice`` (src/ibProcessors3D.hh:202-275,
  src/ibDataExchangeWrappers3D.hh:160-246).
* ``Lattice.diff_forward`` / ``diff_backward`` -- a differentiable window over
  frozen geometry with the hand-written adjoint (no counterpart in the
  reference); ``lbdem_psm.autograd`` wraps it as a ``torch.autograd.Function``.

All compute runs in the HIP library ``lib/liblbdem_psm.so``.  There is no
CPU fallback: if the library is missing, ``load_library`` raises.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence

import numpy as np

Q = 19
OK = 0
ERR_INVALID_ARG = -1
ERR_HIP = -2
ERR_OUT_OF_MEMORY = -3
ERR_CAPACITY = -4
ERR_NO_DEVICE = -5
ERR_STATE = -6
ERR_NUMERIC = -7

FP64 = 0
FP32_DEV = 1
FP16_STORE = 2
PRECISIONS = {"fp64": FP64, "fp32_dev": FP32_DEV, "fp16_store": FP16_STORE}

VOXEL_SUBCELL = 0
VOXEL_OVERLAP = 1
VOXEL_MODES = {"subcell": VOXEL_SUBCELL, "overlap": VOXEL_OVERLAP}

INIT_VEL = 1
AUTO_IMAGES = 2
STEP_ADVANCE = 1
STEP_FROZEN = 2

RADIUS_CONST = 0
RADIUS_UNIFORM = 1
RADIUS_LOGNORMAL = 2

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "liblbdem_psm.so")

# D3Q19 in the Palabos order (c_{i+9} = -c_i) and weights, for host-side
# post-processing (macroscopic fields) only.
C = np.array([
    [0, 0, 0], [-1, 0, 0], [0, -1, 0], [0, 0, -1], [-1, -1, 0], [-1, 1, 0], [-1, 0, -1],
    [-1, 0, 1], [0, -1, -1], [0, -1, 1], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 0],
    [1, -1, 0], [1, 0, 1], [1, 0, -1], [0, 1, 1], [0, 1, -1]], dtype=np.int64)
W = np.array([1 / 3] + [1 / 18] * 3 + [1 / 36] * 6 + [1 / 18] * 3 + [1 / 36] * 6)


class PSMError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"lbdem_psm error {code}: {msg}")
        self.code = code


class Desc(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("periodic", ctypes.c_int32 * 3),
        ("tau", ctypes.c_double),
        ("precision", ctypes.c_int32),
        ("slices_per_dim", ctypes.c_int32),
        ("body_accel", ctypes.c_double * 3),
        ("device", ctypes.c_int32),
        ("voxel_mode", ctypes.c_int32),
        ("max_solid_cells", ctypes.c_int64),
    ]


class WorkloadDesc(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int32 * 3),
        ("periodic", ctypes.c_int32 * 3),
        ("count", ctypes.c_int64),
        ("radius_law", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("a", ctypes.c_double), ("b", ctypes.c_double),
        ("lo", ctypes.c_double), ("hi", ctypes.c_double),
        ("gap", ctypes.c_double), ("wall_gap", ctypes.c_double),
        ("sigma_v", ctypes.c_double), ("sigma_w", ctypes.c_double),
        ("seed", ctypes.c_uint64),
        ("max_attempts", ctypes.c_int64),
    ]


_lib: Optional[ctypes.CDLL] = None

_P = ctypes.c_void_p
_SIGS = {
    "lbdem_psm_version": ([], ctypes.c_char_p),
    "lbdem_psm_desc_init": ([_P], ctypes.c_int),
    "lbdem_psm_check_desc": ([_P], ctypes.c_int),
    "lbdem_psm_check_spheres": ([_P, ctypes.c_int64, _P, _P, _P, _P, _P], ctypes.c_int),
    "lbdem_psm_create": ([_P, ctypes.POINTER(ctypes.c_void_p)], ctypes.c_int),
    "lbdem_psm_destroy": ([_P], ctypes.c_int),
    "lbdem_psm_last_error": ([_P], ctypes.c_char_p),
    "lbdem_psm_init_equilibrium": ([_P, _P, _P], ctypes.c_int),
    "lbdem_psm_set_populations": ([_P, _P], ctypes.c_int),
    "lbdem_psm_get_populations": ([_P, _P], ctypes.c_int),
    "lbdem_psm_set_spheres": ([_P, ctypes.c_int64, _P, _P, _P, _P, _P, _P, _P, ctypes.c_int32,
                               ctypes.c_int32], ctypes.c_int),
    "lbdem_psm_collide_stream": ([_P, ctypes.c_int32], ctypes.c_int),
    "lbdem_psm_sum_force_torque": ([_P, ctypes.c_int64, _P, _P, _P], ctypes.c_int),
    "lbdem_psm_step": ([_P, ctypes.c_int32, ctypes.c_int32], ctypes.c_int),
    "lbdem_psm_get_particle_forces": ([_P, ctypes.c_int64, _P, _P], ctypes.c_int),
    "lbdem_psm_get_particles": ([_P, ctypes.c_int64, _P], ctypes.c_int),
    "lbdem_psm_get_cell_records": ([_P, _P, _P, _P, _P], ctypes.c_int),
    "lbdem_psm_set_cell_records": ([_P, _P, _P, _P], ctypes.c_int),
    "lbdem_psm_subcell_counts": ([_P, ctypes.c_int64, _P, _P, _P, _P, _P], ctypes.c_int),
    "lbdem_psm_subcell_counts_static": ([_P, ctypes.c_int64, _P, _P, _P, _P, _P], ctypes.c_int),
    "lbdem_psm_conservation": ([_P, _P, _P], ctypes.c_int),
    "lbdem_psm_check_health": ([_P, _P, _P], ctypes.c_int),
    "lbdem_psm_stats": ([_P, _P], ctypes.c_int),
    "lbdem_psm_synchronize": ([_P], ctypes.c_int),
    "lbdem_psm_profile": ([_P, ctypes.c_int32], ctypes.c_int),
    "lbdem_psm_profile_read": ([_P, _P, _P, ctypes.c_int32], ctypes.c_int),
    "lbdem_psm_generate_spheres": ([_P, _P, _P, _P, _P, _P], ctypes.c_int),
    "lbdem_psm_set_tau": ([_P, ctypes.c_double], ctypes.c_int),
    "lbdem_psm_clear_records": ([_P], ctypes.c_int),
    "lbdem_psm_perturb_solid_fraction": ([_P, _P], ctypes.c_int),
    "lbdem_psm_diff_forward": ([_P, ctypes.c_int32, ctypes.c_int32, _P], ctypes.c_int),
    "lbdem_psm_diff_backward": ([_P, _P, _P, _P, _P, _P, _P, _P, _P, _P], ctypes.c_int),
    "lbdem_psm_diff_release": ([_P], ctypes.c_int),
    "lbdem_psm_set_periodic_pressure": ([_P, ctypes.c_int32, ctypes.c_int32, ctypes.c_double, ctypes.c_double, _P],
                                        ctypes.c_int),
    "lbdem_psm_init_poiseuille": ([_P, ctypes.c_int32, ctypes.c_double, ctypes.c_double, ctypes.c_double, _P],
                                  ctypes.c_int),
    "lbdem_psm_get_macroscopic": ([_P, _P, _P], ctypes.c_int),
    "lbdem_psm_get_desc": ([_P, _P], ctypes.c_int),
    "lbdem_psm_vti_write": ([ctypes.c_char_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_double,
                             ctypes.c_int32, _P, _P, _P, _P], ctypes.c_int),
    "lbdem_psm_write_vti": ([_P, ctypes.c_char_p, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                             ctypes.c_double], ctypes.c_int),
    "lbdem_psm_checkpoint_write_arrays": ([ctypes.c_char_p, _P, _P, _P, _P, _P, _P], ctypes.c_int),
    "lbdem_psm_checkpoint_read_desc": ([ctypes.c_char_p, _P], ctypes.c_int),
    "lbdem_psm_checkpoint_read_arrays": ([ctypes.c_char_p, _P, _P, _P, _P, _P], ctypes.c_int),
    "lbdem_psm_save_checkpoint": ([_P, ctypes.c_char_p], ctypes.c_int),
    "lbdem_psm_copy_bandwidth": ([_P, ctypes.c_int32, _P, _P], ctypes.c_int),
    "lbdem_psm_load_checkpoint": ([_P, ctypes.c_char_p], ctypes.c_int),
}


def load_library(path: Optional[str] = None) -> ctypes.CDLL:
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise PSMError(ERR_STATE, f"HIP library not built: {p} (run __graft_entry__.build())")
    lib = ctypes.CDLL(p)
    for name, (args, res) in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    if path is None:
        _lib = lib
    return lib


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _f64(a, shape=None) -> np.ndarray:
    out = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        out = out.reshape(shape)
    return out


def _check(rc: int, handle=None):
    if rc != OK:
        lib = load_library()
        msg = lib.lbdem_psm_last_error(handle)
        raise PSMError(rc, msg.decode() if msg else "")


def make_desc(nx, ny, nz, periodic=(1, 1, 1), tau=1.0, precision="fp64", slices_per_dim=5,
              body_accel=(0.0, 0.0, 0.0), device=0, max_solid_cells=0, voxel_mode="subcell") -> Desc:
    d = Desc()
    load_library().lbdem_psm_desc_init(ctypes.byref(d))
    d.nx, d.ny, d.nz = int(nx), int(ny), int(nz)
    for k in range(3):
        d.periodic[k] = int(bool(periodic[k]))
        d.body_accel[k] = float(body_accel[k])
    d.tau = float(tau)
    d.precision = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
    d.slices_per_dim = int(slices_per_dim)
    d.device = int(device)
    d.max_solid_cells = int(max_solid_cells)
    d.voxel_mode = VOXEL_MODES[voxel_mode] if isinstance(voxel_mode, str) else int(voxel_mode)
    return d


def check_desc(desc: Desc) -> int:
    return load_library().lbdem_psm_check_desc(ctypes.byref(desc))


def check_spheres(desc: Desc, x, v, w, r, ids) -> int:
    x, v, w = _f64(x, (-1, 3)), _f64(v, (-1, 3)), _f64(w, (-1, 3))
    r = _f64(r).ravel()
    ids = np.ascontiguousarray(ids, dtype=np.int64).ravel()
    return load_library().lbdem_psm_check_spheres(ctypes.byref(desc), len(r), _ptr(x), _ptr(v),
                                                  _ptr(w), _ptr(r), _ptr(ids))


def generate_spheres(box, count, radius_law=RADIUS_CONST, a=1.0, b=0.0, lo=0.0, hi=1e30,
                     periodic=(1, 1, 1), gap=0.0, wall_gap=0.0, sigma_v=0.0, sigma_w=0.0,
                     seed=1, max_attempts=0):
    """Seeded non-overlapping spheres (lbdem_psm_generate_spheres)."""
    d = WorkloadDesc()
    for k in range(3):
        d.n[k] = int(box[k])
        d.periodic[k] = int(bool(periodic[k]))
    d.count = int(count)
    d.radius_law = int(radius_law)
    d.a, d.b, d.lo, d.hi = float(a), float(b), float(lo), float(hi)
    d.gap, d.wall_gap = float(gap), float(wall_gap)
    d.sigma_v, d.sigma_w = float(sigma_v), float(sigma_w)
    d.seed = int(seed)
    d.max_attempts = int(max_attempts)
    x = np.zeros((count, 3))
    v = np.zeros((count, 3))
    w = np.zeros((count, 3))
    r = np.zeros(count)
    ids = np.zeros(count, dtype=np.int64)
    rc = load_library().lbdem_psm_generate_spheres(ctypes.byref(d), _ptr(x), _ptr(v), _ptr(w),
                                                   _ptr(r), _ptr(ids))
    if rc != 0:
        raise PSMError(ERR_INVALID_ARG, "could not generate the sphere packing")
    return x, v, w, r, ids


def vti_write(path, n, fields, spacing=1.0):
    """Write {name: array} (scalar (Nc,) or vector (3, Nc), host layout) as VTK ImageData."""
    names, ncomp, data, keep = [], [], [], []
    for name, arr in fields.items():
        a = np.ascontiguousarray(arr, dtype=np.float64)
        m = 1 if a.ndim == 1 else a.shape[0]
        keep.append(a)
        names.append(name.encode())
        ncomp.append(m)
    nf = len(names)
    cnames = (ctypes.c_char_p * max(nf, 1))(*names)
    cncomp = np.ascontiguousarray(ncomp, dtype=np.int32)
    cdata = (ctypes.c_void_p * max(nf, 1))(*[a.ctypes.data for a in keep])
    rc = load_library().lbdem_psm_vti_write(str(path).encode(), int(n[0]), int(n[1]), int(n[2]), float(spacing),
                                            nf, ctypes.cast(cnames, ctypes.c_void_p), _ptr(cncomp),
                                            ctypes.cast(cdata, ctypes.c_void_p), None)
    _check(rc)


def checkpoint_write(path, desc, f, sf, ids, us, F):
    """Checkpoint file from host arrays (lbdem_psm_checkpoint_write_arrays)."""
    f = _f64(f).reshape(Q, -1)
    sf, us, F = _f64(sf).ravel(), _f64(us).reshape(3, -1), _f64(F).reshape(3, -1)
    ids = np.ascontiguousarray(ids, dtype=np.int64).ravel()
    _check(load_library().lbdem_psm_checkpoint_write_arrays(str(path).encode(), ctypes.byref(desc), _ptr(f),
                                                            _ptr(sf), _ptr(ids), _ptr(us), _ptr(F)))


def checkpoint_read(path):
    """(desc, dict of arrays) of a checkpoint file."""
    d = Desc()
    lib = load_library()
    _check(lib.lbdem_psm_checkpoint_read_desc(str(path).encode(), ctypes.byref(d)))
    nc = d.nx * d.ny * d.nz
    out = {"f": np.zeros((Q, nc)), "sf": np.zeros(nc), "id": np.zeros(nc, dtype=np.int64),
           "us": np.zeros((3, nc)), "F": np.zeros((3, nc))}
    _check(lib.lbdem_psm_checkpoint_read_arrays(str(path).encode(), _ptr(out["f"]), _ptr(out["sf"]),
                                                _ptr(out["id"]), _ptr(out["us"]), _ptr(out["F"])))
    return d, out


def advance_positions(x, v, n, periodic):
    """Host restatement of the prescribed kinematics kernel: x += v, wrapped
    into [-1/2, n - 1/2) on periodic axes (same operations, bit-identical)."""
    x = np.array(x, dtype=np.float64) + np.asarray(v, dtype=np.float64)
    for d in range(3):
        if periodic[d]:
            L = float(n[d])
            col = x[:, d]
            hi = col >= L - 0.5
            col[hi] = col[hi] - L
            lo = col < -0.5
            col[lo] = col[lo] + L
    return x


class Lattice:
    """A D3Q19 PSM lattice on one GPU (one lbdem_psm handle)."""

    def __init__(self, nx, ny, nz, periodic=(1, 1, 1), tau=1.0, precision="fp64",
                 slices_per_dim=5, body_accel=(0.0, 0.0, 0.0), device=0, max_solid_cells=0,
                 voxel_mode="subcell"):
        self.lib = load_library()
        self.desc = make_desc(nx, ny, nz, periodic, tau, precision, slices_per_dim, body_accel,
                              device, max_solid_cells, voxel_mode)
        self.n = (int(nx), int(ny), int(nz))
        self.nc = self.n[0] * self.n[1] * self.n[2]
        self.periodic = tuple(int(bool(p)) for p in periodic)
        h = ctypes.c_void_p()
        _check(self.lib.lbdem_psm_create(ctypes.byref(self.desc), ctypes.byref(h)))
        self.h = h
        self.np_ = 0

    def close(self):
        if getattr(self, "h", None):
            self.lib.lbdem_psm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _c(self, rc):
        _check(rc, self.h)

    # populations -------------------------------------------------------
    def init_equilibrium(self, rho=None, u=None):
        rho = None if rho is None else _f64(rho).ravel()
        u = None if u is None else _f64(u).reshape(3, -1)
        self._c(self.lib.lbdem_psm_init_equilibrium(self.h, _ptr(rho), _ptr(u)))

    def set_populations(self, f):
        f = _f64(f).reshape(Q, -1)
        assert f.shape[1] == self.nc
        self._c(self.lib.lbdem_psm_set_populations(self.h, _ptr(f)))

    def get_populations(self) -> np.ndarray:
        f = np.empty((Q, self.nc))
        self._c(self.lib.lbdem_psm_get_populations(self.h, _ptr(f)))
        return f

    # coupling ----------------------------------------------------------
    def set_spheres(self, x, v, w, r, ids, types=None, exclude=(), init_vel=False,
                    auto_images=True):
        x, v, w = _f64(x, (-1, 3)), _f64(v, (-1, 3)), _f64(w, (-1, 3))
        r = _f64(r).ravel()
        ids = np.ascontiguousarray(ids, dtype=np.int64).ravel()
        n = len(r)
        t = None if types is None else np.ascontiguousarray(types, dtype=np.int32).ravel()
        ex = np.ascontiguousarray(list(exclude), dtype=np.int32)
        flags = (INIT_VEL if init_vel else 0) | (AUTO_IMAGES if auto_images else 0)
        self._c(self.lib.lbdem_psm_set_spheres(self.h, n, _ptr(x), _ptr(v), _ptr(w), _ptr(r),
                                               _ptr(ids), _ptr(t), _ptr(ex) if len(ex) else None,
                                               len(ex), flags))
        self.np_ = n

    def collide_stream(self, nsteps=1):
        self._c(self.lib.lbdem_psm_collide_stream(self.h, int(nsteps)))

    def sum_force_torque(self, x=None):
        n = self.np_
        F = np.zeros((n, 3))
        T = np.zeros((n, 3))
        xx = None if x is None else _f64(x, (-1, 3))
        self._c(self.lib.lbdem_psm_sum_force_torque(self.h, n, _ptr(xx), _ptr(F), _ptr(T)))
        return F, T

    def step(self, nsteps=1, advance=True, frozen=False):
        """nsteps coupled steps on resident data; frozen: no voxelization
        (and no advance), the records of the last one stay."""
        flags = STEP_FROZEN if frozen else (STEP_ADVANCE if advance else 0)
        self._c(self.lib.lbdem_psm_step(self.h, int(nsteps), flags))

    def particle_forces(self):
        n = self.np_
        F = np.zeros((n, 3))
        T = np.zeros((n, 3))
        self._c(self.lib.lbdem_psm_get_particle_forces(self.h, n, _ptr(F), _ptr(T)))
        return F, T

    def particles(self):
        x = np.zeros((self.np_, 3))
        self._c(self.lib.lbdem_psm_get_particles(self.h, self.np_, _ptr(x)))
        return x

    def cell_records(self):
        sf = np.zeros(self.nc)
        ids = np.zeros(self.nc, dtype=np.int64)
        us = np.zeros((3, self.nc))
        F = np.zeros((3, self.nc))
        self._c(self.lib.lbdem_psm_get_cell_records(self.h, _ptr(sf), _ptr(ids), _ptr(us), _ptr(F)))
        return {"sf": sf, "id": ids, "us": us, "F": F}

    def set_cell_records(self, sf, ids=None, us=None):
        sf = _f64(sf).ravel()
        ids = None if ids is None else np.ascontiguousarray(ids, dtype=np.int64).ravel()
        us = None if us is None else _f64(us).reshape(3, -1)
        self._c(self.lib.lbdem_psm_set_cell_records(self.h, _ptr(sf), _ptr(ids), _ptr(us)))

    def subcell_counts(self, d, r):
        d = _f64(d, (-1, 3))
        r = _f64(r).ravel()
        n = len(r)
        sf = np.zeros(n)
        cnt = np.zeros(n, dtype=np.int32)
        br = np.zeros(n, dtype=np.int32)
        self._c(self.lib.lbdem_psm_subcell_counts(self.h, n, _ptr(d), _ptr(r), _ptr(sf), _ptr(cnt),
                                                  _ptr(br)))
        return sf, cnt, br

    # fields and files --------------------------------------------------
    def macroscopic(self):
        """(rho (Nc,), u (3, Nc)) computed on the device."""
        rho = np.zeros(self.nc)
        u = np.zeros((3, self.nc))
        self._c(self.lib.lbdem_psm_get_macroscopic(self.h, _ptr(rho), _ptr(u)))
        return rho, u

    def write_vti(self, path, spacing=1.0, vel_scale=1.0, rho_scale=1.0, p_scale=1.0):
        self._c(self.lib.lbdem_psm_write_vti(self.h, str(path).encode(), float(spacing), float(vel_scale),
                                             float(rho_scale), float(p_scale)))

    def save_checkpoint(self, path):
        self._c(self.lib.lbdem_psm_save_checkpoint(self.h, str(path).encode()))

    def load_checkpoint(self, path):
        self._c(self.lib.lbdem_psm_load_checkpoint(self.h, str(path).encode()))

    # boundary conditions and initial fields ----------------------------
    def set_periodic_pressure(self, dimension, rho_in, rho_out, range4=None, enable=True):
        """Zhang-Kwok periodic pressure across a periodic axis (lbdem_psm_set_periodic_pressure)."""
        r = None if range4 is None else np.ascontiguousarray(range4, dtype=np.int32)
        self._c(self.lib.lbdem_psm_set_periodic_pressure(self.h, int(bool(enable)), int(dimension),
                                                         float(rho_in), float(rho_out), _ptr(r)))

    def init_poiseuille(self, dimension, rho_hi, rho_lo, u_max, n3=None):
        """PoiseuilleProfileAndPressureGradient at equilibrium (lbdem_psm_init_poiseuille)."""
        n = None if n3 is None else np.ascontiguousarray(n3, dtype=np.int32)
        self._c(self.lib.lbdem_psm_init_poiseuille(self.h, int(dimension), float(rho_hi), float(rho_lo),
                                                   float(u_max), _ptr(n)))

    # differentiable window ---------------------------------------------
    def set_tau(self, tau):
        self._c(self.lib.lbdem_psm_set_tau(self.h, float(tau)))
        self.desc.tau = float(tau)

    def clear_records(self):
        self._c(self.lib.lbdem_psm_clear_records(self.h))

    def perturb_solid_fraction(self, dsf):
        dsf = _f64(dsf).ravel()
        assert dsf.shape[0] == self.nc
        self._c(self.lib.lbdem_psm_perturb_solid_fraction(self.h, _ptr(dsf)))

    # the last differentiable window: (token, nsteps, particles); every
    # diff_forward gets a new token, so a backward pass can tell whether the
    # lattice still holds its window
    _window = None
    _window_token = 0

    def diff_forward(self, nsteps, checkpoint_every):
        """Runs the window; returns (nsteps, n, 6) forces and torques."""
        self._window = None
        ft = np.zeros((int(nsteps), self.np_, 6))
        self._c(self.lib.lbdem_psm_diff_forward(self.h, int(nsteps), int(checkpoint_every),
                                                _ptr(ft)))
        self._window_token += 1
        self._window = (self._window_token, int(nsteps), self.np_)
        return ft

    def window_token(self):
        """Token of the window the lattice holds (None without one)."""
        return None if self._window is None else self._window[0]

    def diff_backward(self, g_ft=None, g_fT=None, want_f0=False, want_cells=True, token=None):
        """Adjoint of the last diff_forward: dict of gradients.  want_cells=False
        skips the dense per-cell sf and B gradients (2 Nc doubles).  token:
        the window_token() taken after that forward; raises if the lattice
        has run another window since.  g_ft must be (nsteps, n, 6) and g_fT
        19 x Nc."""
        if self._window is None:
            raise PSMError(ERR_STATE, "diff_backward without a diff_forward window")
        tok, nsteps, n = self._window
        if token is not None and token != tok:
            raise PSMError(ERR_STATE, "the lattice ran another window after this one's forward pass")
        gft = None if g_ft is None else _f64(g_ft)
        gfT = None if g_fT is None else _f64(g_fT)
        if gft is not None and gft.shape != (nsteps, n, 6):
            raise PSMError(ERR_INVALID_ARG, f"g_ft has shape {gft.shape}, the window needs {(nsteps, n, 6)}")
        if gfT is not None:
            if gfT.size != Q * self.nc:
                raise PSMError(ERR_INVALID_ARG, f"g_fT has {gfT.size} values, the window needs {Q * self.nc}")
            gfT = gfT.reshape(Q, -1)
        gx = np.zeros((n, 3))
        gv = np.zeros((n, 3))
        gw = np.zeros((n, 3))
        gtau = ctypes.c_double()
        gsf = np.zeros(self.nc) if want_cells else None
        gB = np.zeros(self.nc) if want_cells else None
        gf0 = np.zeros((Q, self.nc)) if want_f0 else None
        self._c(self.lib.lbdem_psm_diff_backward(self.h, _ptr(gft), _ptr(gfT), _ptr(gx), _ptr(gv),
                                                 _ptr(gw), ctypes.byref(gtau), _ptr(gsf), _ptr(gB),
                                                 _ptr(gf0)))
        out = {"x": gx, "v": gv, "w": gw, "tau": gtau.value, "sf": gsf, "B": gB}
        if want_f0:
            out["f0"] = gf0
        return out

    def diff_release(self):
        self._window = None
        self._c(self.lib.lbdem_psm_diff_release(self.h))

    # diagnostics -------------------------------------------------------
    def conservation(self):
        m = ctypes.c_double()
        mom = np.zeros(3)
        self._c(self.lib.lbdem_psm_conservation(self.h, ctypes.byref(m), _ptr(mom)))
        return m.value, mom

    def check_health(self):
        mu = ctypes.c_double()
        nf = ctypes.c_int64()
        rc = self.lib.lbdem_psm_check_health(self.h, ctypes.byref(mu), ctypes.byref(nf))
        return rc, mu.value, nf.value

    def stats(self):
        out = np.zeros(4, dtype=np.int64)
        self._c(self.lib.lbdem_psm_stats(self.h, _ptr(out)))
        return {"records": int(out[0]), "entries": int(out[1]), "images": int(out[2]),
                "orphans": int(out[3])}

    def synchronize(self):
        self._c(self.lib.lbdem_psm_synchronize(self.h))

    def copy_bandwidth(self, reps=10):
        """Device copy ceiling of this lattice's population store: (ms per
        copy, bytes read + written per copy), HIP events on the handle's
        stream."""
        ms, nb = ctypes.c_double(), ctypes.c_double()
        self._c(self.lib.lbdem_psm_copy_bandwidth(self.h, int(reps), ctypes.byref(ms), ctypes.byref(nb)))
        return ms.value, nb.value

    def profile(self, enable=True):
        self._c(self.lib.lbdem_psm_profile(self.h, int(bool(enable))))

    def profile_read(self, reset=True):
        ms = np.zeros(4)
        cnt = np.zeros(4, dtype=np.int64)
        self._c(self.lib.lbdem_psm_profile_read(self.h, _ptr(ms), _ptr(cnt), int(bool(reset))))
        names = ("collide_stream", "voxelize", "force_sum", "advance")
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(names)}


def macroscopic(f: np.ndarray, body_accel=(0.0, 0.0, 0.0)):
    """Density and velocity from f - t populations, shape (19, Nc)."""
    f = np.asarray(f, dtype=np.float64)
    rho_bar = f.sum(axis=0)
    j = C.T.astype(np.float64) @ f
    rho = 1.0 + rho_bar
    u = j / rho + 0.5 * np.asarray(body_accel, dtype=np.float64)[:, None]
    return rho, u
