"""ctypes binding of the CPU oracle (oracle/liboracle_psm.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
the cpu_baseline leg of bench.py, never by the product.  Method names and
array conventions follow lbdem_psm.Lattice so that parity tests read alike.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_psm.so")
Q = 19
_P = ctypes.c_void_p
_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"oracle not built: {LIB_PATH} (make -C oracle)")
    lib = ctypes.CDLL(LIB_PATH)
    sigs = {
        "orc_create": ([ctypes.c_int, ctypes.c_int, ctypes.c_int, _P, ctypes.c_double, ctypes.c_int, _P], _P),
        "orc_destroy": ([_P], None),
        "orc_init_equilibrium": ([_P, _P, _P], None),
        "orc_set_populations": ([_P, _P], None),
        "orc_get_populations": ([_P, _P], None),
        "orc_set_spheres": ([_P, ctypes.c_int64, _P, _P, _P, _P, _P, _P, _P, ctypes.c_int, ctypes.c_int,
                             ctypes.c_int], ctypes.c_int),
        "orc_collide_stream": ([_P, ctypes.c_int, ctypes.c_int], None),
        "orc_sum_force_torque": ([_P, ctypes.c_int64, _P, _P, _P], ctypes.c_int),
        "orc_get_cell_records": ([_P, _P, _P, _P, _P], None),
        "orc_set_cell_records": ([_P, _P, _P, _P], None),
        "orc_calc_solid_fraction": ([ctypes.c_double] * 4 + [ctypes.c_int, _P, _P], ctypes.c_double),
        "orc_conservation": ([_P, _P, _P], None),
        "orc_equilibrium": ([ctypes.c_double, _P, _P], None),
        "orc_collide_cell": ([_P, ctypes.c_double, ctypes.c_double, _P, _P], None),
        "orc_last_n_voxelized": ([_P], ctypes.c_int64),
        "orc_overlap_fraction": ([ctypes.c_double, ctypes.c_double, _P, _P], ctypes.c_double),
        "orc_set_voxel_mode": ([_P, ctypes.c_int], None),
        "orc_set_threads": ([_P, ctypes.c_int], None),
        "orc_set_periodic_pressure": ([_P, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, _P], None),
        "orc_init_poiseuille": ([_P, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, _P], None),
    }
    for name, (a, r) in sigs.items():
        fn = getattr(lib, name)
        fn.argtypes = a
        fn.restype = r
    _lib = lib
    return lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def calc_solid_fraction(dx, dy, dz, r, N):
    n = ctypes.c_int32()
    br = ctypes.c_int32()
    sf = load().orc_calc_solid_fraction(float(dx), float(dy), float(dz), float(r), int(N),
                                        ctypes.byref(n), ctypes.byref(br))
    return sf, n.value, br.value


def overlap_fraction(d, r):
    """Analytic sphere-cell overlap: (sf, dsf/dd, branch)."""
    der = ctypes.c_double()
    br = ctypes.c_int32()
    sf = load().orc_overlap_fraction(float(d), float(r), ctypes.byref(der), ctypes.byref(br))
    return sf, der.value, br.value


def equilibrium(rho_bar, j):
    j = np.ascontiguousarray(j, dtype=np.float64)
    out = np.zeros(Q)
    load().orc_equilibrium(float(rho_bar), _p(j), _p(out))
    return out


def collide_cell(f, tau, sf, us):
    f = np.array(f, dtype=np.float64)
    us = np.ascontiguousarray(us, dtype=np.float64)
    F = np.zeros(3)
    load().orc_collide_cell(_p(f), float(tau), float(sf), _p(us), _p(F))
    return f, F


class OracleLattice:
    def __init__(self, nx, ny, nz, periodic=(1, 1, 1), tau=1.0, slices_per_dim=5,
                 body_accel=(0.0, 0.0, 0.0), voxel_mode=0):
        self.lib = load()
        self.n = (int(nx), int(ny), int(nz))
        self.nc = self.n[0] * self.n[1] * self.n[2]
        per = np.ascontiguousarray([int(bool(p)) for p in periodic], dtype=np.int32)
        acc = np.ascontiguousarray(body_accel, dtype=np.float64)
        self.h = self.lib.orc_create(self.n[0], self.n[1], self.n[2], _p(per), float(tau),
                                     int(slices_per_dim), _p(acc))
        if not self.h:
            raise ValueError("invalid oracle lattice parameters")
        self.lib.orc_set_voxel_mode(self.h, int(voxel_mode))
        self.np_ = 0

    def close(self):
        if self.h:
            self.lib.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init_equilibrium(self, rho=None, u=None):
        rho = None if rho is None else np.ascontiguousarray(rho, dtype=np.float64).ravel()
        u = None if u is None else np.ascontiguousarray(u, dtype=np.float64).reshape(3, -1)
        self.lib.orc_init_equilibrium(self.h, _p(rho), _p(u))

    def set_populations(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64).reshape(Q, -1)
        self.lib.orc_set_populations(self.h, _p(f))

    def get_populations(self):
        f = np.empty((Q, self.nc))
        self.lib.orc_get_populations(self.h, _p(f))
        return f

    def set_spheres(self, x, v, w, r, ids, types=None, exclude=(), init_vel=False,
                    auto_images=True):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
        v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1, 3)
        w = np.ascontiguousarray(w, dtype=np.float64).reshape(-1, 3)
        r = np.ascontiguousarray(r, dtype=np.float64).ravel()
        ids = np.ascontiguousarray(ids, dtype=np.int64).ravel()
        t = None if types is None else np.ascontiguousarray(types, dtype=np.int32).ravel()
        ex = np.ascontiguousarray(list(exclude), dtype=np.int32)
        rc = self.lib.orc_set_spheres(self.h, len(r), _p(x), _p(v), _p(w), _p(r), _p(ids), _p(t),
                                      _p(ex) if len(ex) else None, len(ex), int(bool(init_vel)),
                                      int(bool(auto_images)))
        if rc != 0:
            raise ValueError(f"orc_set_spheres failed: {rc}")
        self.np_ = len(r)

    def set_threads(self, nthreads):
        """x-slab threads of set_spheres and sum_force_torque (1 = sequential)."""
        self.lib.orc_set_threads(self.h, int(nthreads))

    def collide_stream(self, nsteps=1, nthreads=0):
        self.lib.orc_collide_stream(self.h, int(nsteps), int(nthreads))

    def sum_force_torque(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
        n = len(x)
        F = np.zeros((n, 3))
        T = np.zeros((n, 3))
        self.lib.orc_sum_force_torque(self.h, n, _p(x), _p(F), _p(T))
        return F, T

    def cell_records(self):
        sf = np.zeros(self.nc)
        ids = np.zeros(self.nc, dtype=np.int64)
        us = np.zeros((3, self.nc))
        F = np.zeros((3, self.nc))
        self.lib.orc_get_cell_records(self.h, _p(sf), _p(ids), _p(us), _p(F))
        return {"sf": sf, "id": ids, "us": us, "F": F}

    def set_cell_records(self, sf, ids=None, us=None):
        sf = np.ascontiguousarray(sf, dtype=np.float64).ravel()
        ids = None if ids is None else np.ascontiguousarray(ids, dtype=np.int64).ravel()
        us = None if us is None else np.ascontiguousarray(us, dtype=np.float64).reshape(3, -1)
        self.lib.orc_set_cell_records(self.h, _p(sf), _p(ids), _p(us))

    def set_periodic_pressure(self, dimension, rho_in, rho_out, range4=None, enable=True):
        r = None if range4 is None else np.ascontiguousarray(range4, dtype=np.int32)
        self.lib.orc_set_periodic_pressure(self.h, int(bool(enable)), int(dimension), float(rho_in),
                                           float(rho_out), _p(r))

    def init_poiseuille(self, dimension, rho_hi, rho_lo, u_max, n3=None):
        n = None if n3 is None else np.ascontiguousarray(n3, dtype=np.int32)
        self.lib.orc_init_poiseuille(self.h, int(dimension), float(rho_hi), float(rho_lo), float(u_max), _p(n))

    def conservation(self):
        m = ctypes.c_double()
        mom = np.zeros(3)
        self.lib.orc_conservation(self.h, ctypes.byref(m), _p(mom))
        return m.value, mom

    def n_voxelized(self):
        return int(self.lib.orc_last_n_voxelized(self.h))
