"""Benchmark of the MI355X PSM hot path (D3Q19 Noble-Torczynski collide+stream,
sphere voxelization, per-particle force/torque sum) on one GPU.

Headline workload: the production-scale rectangular channel -- 512x512x768
D3Q19, walls in x/y, pressure-driven along z (Zhang-Kwok periodic pressure
from a Poiseuille initial field), tau 0.55, 100000 non-overlapping spheres
with radii U[3,6] filling ~30 % of the volume, 8^3 sub-cell sampling,
prescribed kinematics (x += v each step), FP64.  One step = voxelize + fused
collide/stream + force/torque sum, all resident in HBM.  `value` is the
whole-step MLUPS of the FP64 run, the median over fresh lattices (each
lattice's memory placement sets its speed for its lifetime, +-8 % between
placements).  The FP32-deviation and FP16-storage runs of the same workload
are reported beside it, each with its own roofline, and with drag, torque
and velocity-field errors against the FP64 run of the same seed after the
configuration's 200 steps (norm-wise and per particle).  Also reported: the
512^3 periodic box at the same sphere density, the 256^3 dense suspension,
the 384^3 polydisperse suspension over its 1000 steps, the differentiable
128x128x256 channel and the adjoint of the last 50 steps of the channel,
the host-driven exchange loop (set_spheres -> collide_stream ->
sum_force_torque through the C ABI) beside the resident step, and the
device copy ceiling of the population store.

cpu_baseline: the CPU restatement of the reference algorithm on a slab of
the headline channel (512x512x96 with its sphere density), all host threads
(x-slabs, as the reference's MPI ranks), plus one step on one core.

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "lbdemcoupling-public_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

import lbdem_psm as L  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "MLUPS of PSM step (FP64 vs FP32-dev) + %HBM roofline; drag rel-err vs FP64"
BYTES_PER_POP = {"fp64": 8, "fp32_dev": 4, "fp16_store": 2}
PRECISIONS = ("fp64", "fp32_dev", "fp16_store")
KERNEL = {"fp64": "k_collide_stream<FP64>", "fp32_dev": "k_collide_stream<FP32_DEV>",
          "fp16_store": "k_collide_stream<FP16_STORE>"}

WORKLOADS = {
    "channel512": dict(n=(512, 512, 768), periodic=(0, 0, 1), tau=0.55, count=100000,
                       law=L.RADIUS_UNIFORM, a=3.0, b=6.0, lo=0.0, hi=0.0, body=(0.0, 0.0, 0.0),
                       pressure=dict(dim=2, rho_hi=1.0, rho_lo=1.0 - 1e-4, u_max=0.05),
                       # slow prescribed kinematics (0.4 lu in 200 steps) keep the spheres apart
                       # and off the walls: with v ~ N(0, 0.01) they pass through each other and
                       # the walls within 200 steps and the flow leaves the stable range
                       sigma_v=0.002, sigma_w=2e-4, wall_gap=2.0, full_steps=200,
                       desc="channel 512x512x768, walls x/y, pressure-driven along z (Zhang-Kwok "
                            "periodic pressure rho 1 -> 1-1e-4, Poiseuille initial field u_max 0.05), "
                            "tau 0.55, 100000 spheres r~U[3,6], v~N(0,0.002), w~N(0,2e-4)"),
    # one eighth of the channel cut along x, the axis the population store's
    # rows are ordered by (row = x * ny + y): one GPU's share of an 8-way
    # split, a contiguous range of rows
    "channel512_slab": dict(n=(64, 512, 768), periodic=(0, 0, 1), tau=0.55, count=12500,
                            law=L.RADIUS_UNIFORM, a=3.0, b=6.0, lo=0.0, hi=0.0, body=(0.0, 0.0, 0.0),
                            pressure=dict(dim=2, rho_hi=1.0, rho_lo=1.0 - 1e-4, u_max=0.05),
                            sigma_v=0.002, sigma_w=2e-4, wall_gap=2.0,
                            desc="slab 64x512x768 of the channel (1/8 of it along x, same sphere "
                                 "density and pressure gradient): 12500 spheres r~U[3,6]"),
    "cube512": dict(n=(512, 512, 512), periodic=(1, 1, 1), tau=0.55, count=66667,
                    law=L.RADIUS_UNIFORM, a=3.0, b=6.0, lo=0.0, hi=0.0, body=(0.0, 0.0, 0.0),
                    sigma_v=0.002, sigma_w=2e-4, full_steps=200,
                    desc="periodic 512^3, tau 0.55, 66667 spheres r~U[3,6] (the channel's sphere "
                         "density), v~N(0,0.002), w~N(0,2e-4)"),
    "dense256": dict(n=(256, 256, 256), periodic=(1, 1, 1), tau=0.7, count=2000,
                     law=L.RADIUS_UNIFORM, a=4.0, b=8.0, lo=0.0, hi=0.0, body=(0.0, 0.0, 0.0),
                     desc="periodic dense suspension 256^3, tau 0.7, 2000 spheres r~U[4,8], "
                          "v~N(0,0.01), w~N(0,1e-3)"),
    "poly384": dict(n=(384, 384, 384), periodic=(1, 1, 1), tau=0.6, count=20000,
                    law=L.RADIUS_LOGNORMAL, a=5.0, b=0.25, lo=3.0, hi=10.0, body=(0.0, 0.0, 0.0),
                    # over its 1000 steps the configuration's v ~ N(0, 0.01) carries the
                    # spheres through each other and the flow past |u| = 0.3 after ~200
                    # steps (FP64 as well); slower kinematics (1 lu in 1000 steps) keep it
                    # stable (the packing leaves no room for a gap between the spheres)
                    sigma_v=0.001, sigma_w=1e-4, full_steps=1000,
                    desc="periodic polydisperse 384^3, tau 0.6, 20000 spheres lognormal(5,0.25) in "
                         "[3,10], v~N(0,0.001), w~N(0,1e-4), 1000 steps"),
}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_spheres(wl, seed=2024):
    return L.generate_spheres(wl["n"], wl["count"], wl["law"], a=wl["a"], b=wl["b"], lo=wl["lo"],
                              hi=wl["hi"] if wl["hi"] > 0 else 1e30, periodic=wl["periodic"],
                              gap=wl.get("gap", 0.0), wall_gap=wl.get("wall_gap", 1.0),
                              sigma_v=wl.get("sigma_v", 0.01), sigma_w=wl.get("sigma_w", 1e-3),
                              seed=seed, max_attempts=10 ** 8)


def new_lattice(wl, precision, voxel_mode="subcell"):
    lat = L.Lattice(*wl["n"], periodic=wl["periodic"], tau=wl["tau"], precision=precision,
                    slices_per_dim=8, body_accel=wl["body"], voxel_mode=voxel_mode)
    pr = wl.get("pressure")
    if pr:
        lat.init_poiseuille(pr["dim"], pr["rho_hi"], pr["rho_lo"], pr["u_max"])
        lat.set_periodic_pressure(pr["dim"], pr["rho_hi"], pr["rho_lo"])
    return lat


def run_gpu(wl, precision, steps, warmup, spheres, full_steps=0, copy_reps=0):
    """warmup + `steps` timed coupled steps on a fresh lattice.  full_steps >
    0: continue to that many steps in all, then read the particle forces and
    the velocity field (for the error against FP64).  copy_reps > 0: the
    copy ceiling of this lattice's population store afterwards."""
    nc = int(np.prod(wl["n"]))
    x, v, w, r, ids = spheres
    lat = new_lattice(wl, precision)
    try:
        lat.set_spheres(x, v, w, r, ids)
        lat.step(warmup, advance=True)
        lat.synchronize()
        lat.profile(True)
        lat.profile_read(reset=True)
        t0 = time.perf_counter()
        lat.step(steps, advance=True)
        lat.synchronize()
        t1 = time.perf_counter()
        prof = lat.profile_read(reset=True)
        lat.profile(False)
        st = lat.stats()
        out = {}
        if full_steps > 0:
            rest = full_steps - warmup - steps
            if rest > 0:
                lat.step(rest, advance=True)
            F, T = lat.particle_forces()
            rc, umax, _ = lat.check_health()
            out.update(F=F, T=T, u=lat.macroscopic()[1], max_u=umax, health_rc=rc, steps_total=max(full_steps, warmup + steps))
        if copy_reps > 0:
            ms, nbytes = lat.copy_bandwidth(copy_reps)
            out["copy_gbs"] = nbytes / (ms * 1e-3) / 1e9
    finally:
        lat.close()
    wall = t1 - t0
    cs_ms, cs_n = prof["collide_stream"]
    cs_avg_s = cs_ms / max(cs_n, 1) * 1e-3
    b_lu = 2 * 19 * BYTES_PER_POP[precision]
    out.update({
        "precision": precision,
        "mlups": nc * steps / wall / 1e6,
        "ms_per_step": wall / steps * 1e3,
        "kernel_ms": cs_avg_s * 1e3,
        "kernel_gbs": b_lu * nc / cs_avg_s / 1e9,
        "voxelize_ms": prof["voxelize"][0] / max(prof["voxelize"][1], 1),
        "force_ms": prof["force_sum"][0] / max(prof["force_sum"][1], 1),
        "advance_ms": prof["advance"][0] / max(prof["advance"][1], 1),
        "bytes_per_lu": b_lu,
        "nc": nc,
        "records": st["records"], "images": st["images"], "orphans": st["orphans"],
    })
    return out


def run_line(wl, precision, steps, warmup, spheres, repeats, full_steps=0, copy_reps=0):
    """`repeats` fresh lattices; the median repetition by step rate, with the
    forces and fields of the first (identical in every repetition)."""
    runs = []
    for k in range(max(1, repeats)):
        runs.append(run_gpu(wl, precision, steps, warmup, spheres, full_steps if k == 0 else 0,
                            copy_reps if k == 0 else 0))
        log(f"  {precision} rep {k + 1}: {runs[-1]['mlups']:.0f} MLUPS step, collide "
            f"{runs[-1]['kernel_ms']:.3f} ms ({runs[-1]['kernel_gbs']:.0f} GB/s), voxelize "
            f"{runs[-1]['voxelize_ms']:.3f} ms, force {runs[-1]['force_ms']:.3f} ms")
    med = dict(sorted(runs, key=lambda rr: rr["mlups"])[len(runs) // 2])
    for key in ("F", "T", "u", "max_u", "health_rc", "copy_gbs", "steps_total"):
        if key in runs[0]:
            med[key] = runs[0][key]
    med["runs_mlups"] = [round(rr["mlups"], 1) for rr in runs]
    med["runs_kernel_ms"] = [round(rr["kernel_ms"], 4) for rr in runs]
    return med


def rel(a, b):
    den = np.abs(b).max()
    return float(np.abs(a - b).max() / den) if den > 0 else 0.0


def particle_err(a, ref):
    """Per-particle relative error |a_p - ref_p| / |ref_p| (vector norms), with
    |ref_p| floored at 1e-3 of the median |ref|, so that particles whose load
    is nearly zero do not divide by round-off."""
    mag = np.linalg.norm(ref, axis=1)
    floor = 1e-3 * float(np.median(mag))
    e = np.linalg.norm(a - ref, axis=1) / np.maximum(mag, max(floor, 1e-300))
    return {"median": float(np.median(e)), "p99": float(np.percentile(e, 99)), "max": float(e.max()),
            "floor": floor}


def field_err(u, ref):
    """Relative L2 and Linf error of a velocity field against the FP64 one."""
    if u is None or ref is None:
        return None
    d = u - ref
    return {"l2": float(np.linalg.norm(d) / np.linalg.norm(ref)), "linf": float(np.abs(d).max() / np.abs(ref).max())}


TRAFFIC_JSON = os.path.join(ROOT, "profiles", "collide_stream_traffic.json")


def measured_traffic(workload, kernel):
    """HBM bytes per launch of a kernel from the calibrated FETCH_SIZE /
    WRITE_SIZE counters of that workload, or None."""
    try:
        t = json.load(open(TRAFFIC_JSON))
    except (OSError, ValueError):
        return None
    v = t.get("workloads", {}).get(workload, {}).get("bytes_per_launch", {}).get(kernel)
    return None if v is None else int(v)


def roofline(rr, workload):
    k = KERNEL[rr["precision"]]
    return {"bound": "hbm", "achieved": round(rr["kernel_gbs"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(rr["kernel_gbs"] / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(workload, k),
            "kernel": k,
            "algorithmic_bytes_per_launch": rr["bytes_per_lu"] * rr["nc"],
            "kernel_ms": round(rr["kernel_ms"], 4)}


def summary(rr, ref, workload):
    out = {"mlups": round(rr["mlups"], 1), "ms_per_step": round(rr["ms_per_step"], 4),
           "runs_mlups": rr.get("runs_mlups"), "runs_kernel_ms": rr.get("runs_kernel_ms"),
           "kernel_ms": round(rr["kernel_ms"], 4), "kernel_gbs": round(rr["kernel_gbs"], 1),
           "roofline": roofline(rr, workload),
           "step_roofline_frac": round(rr["mlups"] * 1e6 * rr["bytes_per_lu"] / 1e9 / HBM_PEAK_GBS, 4),
           "voxelize_ms": round(rr["voxelize_ms"], 4), "force_ms": round(rr["force_ms"], 4),
           "advance_ms": round(rr["advance_ms"], 4),
           "records": rr["records"], "images": rr["images"], "orphans": rr["orphans"]}
    if "copy_gbs" in rr:
        out["device_copy_gbs"] = round(rr["copy_gbs"], 1)
    if "F" in rr:
        out["steps_compared"] = rr["steps_total"]
        out["max_u"] = rr["max_u"]
        out["health_rc"] = rr["health_rc"]
        if ref is not None and rr is not ref:
            out["drag_rel_err_vs_fp64"] = rel(rr["F"], ref["F"])
            out["torque_rel_err_vs_fp64"] = rel(rr["T"], ref["T"])
            out["drag_particle_rel_err_vs_fp64"] = particle_err(rr["F"], ref["F"])
            out["torque_particle_rel_err_vs_fp64"] = particle_err(rr["T"], ref["T"])
            out["velocity_err_vs_fp64"] = field_err(rr.get("u"), ref.get("u"))
    return out


def run_workload(key, steps, warmup, repeats, copy_reps=0, spheres=None):
    """All three precisions of one workload; errors against its FP64 run."""
    wl = WORKLOADS[key]
    if spheres is None:
        log(f"{key}: generating {wl['count']} spheres")
        spheres = make_spheres(wl)
    full = wl.get("full_steps", warmup + steps)
    res = {}
    for prec in PRECISIONS:
        log(f"{key} {prec}: {repeats} x ({warmup} warmup + {steps} timed steps), errors after {full} steps")
        res[prec] = run_line(wl, prec, steps, warmup, spheres, repeats, full_steps=full,
                             copy_reps=copy_reps if prec == "fp64" else 0)
    out = {"workload": wl["desc"], "cells": int(np.prod(wl["n"])), "spheres": wl["count"]}
    for prec in PRECISIONS:
        out[prec] = summary(res[prec], res["fp64"], key)
    out["fp32_over_fp64_mlups"] = round(res["fp32_dev"]["mlups"] / res["fp64"]["mlups"], 3)
    out["fp16_over_fp64_mlups"] = round(res["fp16_store"]["mlups"] / res["fp64"]["mlups"], 3)
    return out, res, spheres


ADJOINT_CASE = dict(n=(128, 128, 256), periodic=(0, 0, 1), tau=0.8, body=(0.0, 0.0, 1e-6), count=200,
                    radius=5.0, sigma_v=0.005, steps=500, checkpoint_every=50,
                    desc="differentiable channel 128x128x256, walls x/y, body force 1e-6 along z, "
                         "200 spheres r=5, v~N(0,0.005), analytic overlap, adjoint of 500 steps, "
                         "checkpoint every 50")


def run_adjoint(precision):
    """Forward window and hand-written adjoint of d(total z-drag)/d(params)."""
    c = ADJOINT_CASE
    x, v, w, r, ids = L.generate_spheres(c["n"], c["count"], L.RADIUS_CONST, a=c["radius"],
                                         periodic=c["periodic"], gap=0.5, wall_gap=1.5,
                                         sigma_v=c["sigma_v"], sigma_w=0.0, seed=44)
    lat = L.Lattice(*c["n"], periodic=c["periodic"], tau=c["tau"], precision=precision,
                    body_accel=c["body"], voxel_mode="overlap")
    try:
        lat.set_spheres(x, v, w, r, ids)
        # The first window of a shape also allocates the handle's state pool
        # (10 checkpoints and 49 recompute states); it is timed on its own.
        # The timed window is the next one of the same shape, as in a loop of
        # windows, where the pool is reused.
        lat.synchronize()
        c0 = time.perf_counter()
        ft = lat.diff_forward(c["steps"], c["checkpoint_every"])
        c1 = time.perf_counter()
        G = np.zeros_like(ft)
        G[:, :, 2] = 1.0
        lat.diff_backward(G)
        c2 = time.perf_counter()
        t0 = time.perf_counter()
        ft = lat.diff_forward(c["steps"], c["checkpoint_every"])
        t1 = time.perf_counter()
        g = lat.diff_backward(G)
        t2 = time.perf_counter()
        # the plain forward of the same frozen geometry: collide+stream and the
        # device force sum per step, no checkpoints, no host copies
        lat.step(2, frozen=True)
        lat.synchronize()
        t3 = time.perf_counter()
        lat.step(c["steps"], frozen=True)
        lat.synchronize()
        t4 = time.perf_counter()
        lat.diff_release()
    finally:
        lat.close()
    fwd, bwd, plain = t1 - t0, t2 - t1, t4 - t3
    return {"forward_s": round(fwd, 4), "backward_s": round(bwd, 4), "plain_forward_s": round(plain, 4),
            "backward_over_forward": round(bwd / fwd, 3),
            "forward_over_plain_forward": round(fwd / plain, 3),
            "total_over_plain_forward": round((fwd + bwd) / plain, 3),
            "first_window_forward_s": round(c1 - c0, 4), "first_window_backward_s": round(c2 - c1, 4),
            "drag_z_total": float(ft[:, :, 2].sum()),
            "grad_tau": g["tau"], "grad_v_norm": float(np.linalg.norm(g["v"])),
            "grad_x_norm": float(np.linalg.norm(g["x"])), "grad_w_norm": float(np.linalg.norm(g["w"])),
            "grad_sf_norm": float(np.linalg.norm(g["sf"])), "_g": g}


ADJOINT_C5 = dict(pre_steps=150, window=50, budget=10,
                  desc="the 512x512x768 pressure-driven channel in FP32-dev with the analytic "
                       "overlap: 150 coupled steps, then the adjoint of the last 50 with a budget of 10 "
                       "population states (segments of 10, 9, ... steps, each finished checkpoint reused "
                       "for recompute), d(total z-drag)/d(tau, v, w, x, sf); central differences of the "
                       "FP32 forward along tau and along all particles' v_z and w_z")


def device_memory_gb():
    """(used, total) device memory in GB from hipMemGetInfo."""
    try:
        hip = ctypes.CDLL("libamdhip64.so")
        free, total = ctypes.c_size_t(), ctypes.c_size_t()
        if hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) != 0:
            return None
        return round((total.value - free.value) / 1e9, 1), round(total.value / 1e9, 1)
    except OSError:
        return None


def time_adjoint_c5(spheres, xw, markers=False):
    """The timed part of the 512x512x768 adjoint line: a warm-up window (the
    state pool and adjoint buffers), 150 coupled steps, the 50-step window
    forward and backward (d(total z-drag)), then 50 plain steps of the same
    frozen geometry.  Returns (forward s, backward s, plain s, forces,
    gradients, device memory, state GB).  markers: one untimed copy of the
    population store (k_copy16, known size) before the forward, the backward
    and the plain steps, to split a kernel trace into the three phases."""
    wl, c, pr = WORKLOADS["channel512"], ADJOINT_C5, WORKLOADS["channel512"]["pressure"]
    x, v, w, r, ids = spheres
    lat = new_lattice(wl, "fp32_dev", voxel_mode="overlap")
    try:
        lat.set_spheres(x, v, w, r, ids)
        # warm-up of both passes at the same budget (one segment: all M states):
        # allocates the state pool and the adjoint buffers, which the timed
        # window reuses; then the same start as objective() from a fresh history
        lat.diff_forward(c["budget"], -c["budget"])
        lat.diff_backward(np.zeros((c["budget"], len(r), 6)), want_cells=False)
        lat.clear_records()
        lat.init_poiseuille(pr["dim"], pr["rho_hi"], pr["rho_lo"], pr["u_max"])
        lat.set_spheres(x, v, w, r, ids)
        lat.step(c["pre_steps"], advance=True)
        lat.set_spheres(xw, v, w, r, ids)
        if markers:
            lat.copy_bandwidth(1)
        lat.synchronize()
        t0 = time.perf_counter()
        ft = lat.diff_forward(c["window"], -c["budget"])
        t1 = time.perf_counter()
        G = np.zeros_like(ft)
        G[:, :, 2] = 1.0
        if markers:
            lat.copy_bandwidth(1)
            lat.synchronize()
        t1b = time.perf_counter()
        g = lat.diff_backward(G, want_cells=False)
        t2 = time.perf_counter()
        mem = device_memory_gb()
        lat.diff_release()
        if markers:
            lat.copy_bandwidth(1)
        lat.synchronize()
        t3 = time.perf_counter()
        lat.step(c["window"], frozen=True)
        lat.synchronize()
        t4 = time.perf_counter()
        state_gb = lat.nc * 19 * 4 / 1e9
    finally:
        lat.close()
    return t1 - t0, t2 - t1b, t4 - t3, ft, g, mem, state_gb


def run_adjoint_c5(spheres):
    """Adjoint of the last 50 steps of the 512x512x768 channel (FP32)."""
    wl, c, pr = WORKLOADS["channel512"], ADJOINT_C5, WORKLOADS["channel512"]["pressure"]
    x, v, w, r, ids = spheres
    xw = np.array(x, dtype=np.float64)
    for _ in range(c["pre_steps"]):
        xw = L.advance_positions(xw, v, wl["n"], wl["periodic"])

    def objective(dtau=0.0, dvz=0.0, dwz=0.0):
        lat = new_lattice(wl, "fp32_dev", voxel_mode="overlap")
        try:
            lat.set_spheres(x, v, w, r, ids)
            lat.step(c["pre_steps"], advance=True)
            if dtau:
                lat.set_tau(wl["tau"] + dtau)
            vv, ww = np.array(v, dtype=np.float64), np.array(w, dtype=np.float64)
            vv[:, 2] += dvz
            ww[:, 2] += dwz
            lat.set_spheres(xw, vv, ww, r, ids)
            ft = lat.diff_forward(c["window"], -c["budget"])
        finally:
            lat.close()
        return float(ft[:, :, 2].sum())

    fwd, bwd, plain, ft, g, mem, state_gb = time_adjoint_c5(spheres, xw)
    fd = {}
    for name, kw, eps, pred in (("tau", "dtau", 1e-4, g["tau"]),
                                ("all_v_z", "dvz", 1e-4, float(np.sum(g["v"][:, 2]))),
                                ("all_w_z", "dwz", 1e-4, float(np.sum(g["w"][:, 2])))):
        jp = objective(**{kw: eps})
        jm = objective(**{kw: -eps})
        fdv = (jp - jm) / (2 * eps)
        fd[name] = {"adjoint": pred, "central_difference": fdv, "eps": eps, "J_plus": jp, "J_minus": jm,
                    "rel_err": abs(pred - fdv) / max(abs(fdv), 1e-300)}
    return {"workload": c["desc"], "forward_s": round(fwd, 4), "backward_s": round(bwd, 4),
            "plain_forward_s": round(plain, 4), "backward_over_forward": round(bwd / fwd, 3),
            "forward_over_plain_forward": round(fwd / plain, 3),
            "total_over_plain_forward": round((fwd + bwd) / plain, 3),
            "population_state_gb": round(state_gb, 2), "state_budget": c["budget"],
            "device_memory_used_after_backward_gb": mem[0] if mem else None,
            "device_memory_total_gb": mem[1] if mem else None,
            "drag_z_total": float(ft[:, :, 2].sum()),
            "grad_tau": g["tau"], "grad_v_norm": float(np.linalg.norm(g["v"])),
            "grad_x_norm": float(np.linalg.norm(g["x"])), "grad_w_norm": float(np.linalg.norm(g["w"])),
            "central_differences_fp32": fd,
            "note": "positions are checked by central differences on the smaller windows (FP64): "
                    "moving all 100000 spheres at once carries cells across the reference's "
                    "solid-fraction thresholds (0.001, 0.999), where the objective jumps"}


def run_exchange(key, precision, steps, warmup, spheres):
    """The reference's coupling loop through the C ABI with host particle
    arrays (setSpheresOnLattice -> collideAndStream -> getForcesFromLattice,
    sedimentingSphere.cpp:212-225), against the resident step of the same
    workload."""
    wl = WORKLOADS[key]
    nc = int(np.prod(wl["n"]))
    x, v, w, r, ids = spheres
    lat = new_lattice(wl, precision)
    try:
        xs = np.array(x, dtype=np.float64)

        def one():
            nonlocal xs
            lat.set_spheres(xs, v, w, r, ids)
            lat.collide_stream(1)
            F, T = lat.sum_force_torque(xs)
            xs = L.advance_positions(xs, v, wl["n"], wl["periodic"])
            return F

        for _ in range(warmup):
            one()
        lat.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            one()
        lat.synchronize()
        t1 = time.perf_counter()
    finally:
        lat.close()
    return {"workload": wl["desc"], "precision": precision, "steps": steps,
            "mlups": round(nc * steps / (t1 - t0) / 1e6, 1), "ms_per_step": round((t1 - t0) / steps * 1e3, 4),
            "loop": "host arrays each step: set_spheres (upload, voxelize) -> collide_stream -> "
                    "sum_force_torque (force sum, download, map(tag) fold on the host)"}


def cpu_baseline(nsteps):
    """The CPU restatement of the reference algorithm on a slab of the headline
    channel: every step voxelize + collide/stream + force sum, on all host
    threads (x-slabs, as the reference's MPI ranks), then one step on one core."""
    from oracle import oracle as O
    wl = WORKLOADS["channel512_slab"]
    n = wl["n"]
    nc = int(np.prod(n))
    x, v, w, r, ids = make_spheres(wl)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    o = O.OracleLattice(*n, periodic=wl["periodic"], tau=wl["tau"], slices_per_dim=8, body_accel=wl["body"])
    pr = wl["pressure"]
    o.init_poiseuille(pr["dim"], pr["rho_hi"], pr["rho_lo"], pr["u_max"])
    o.set_periodic_pressure(pr["dim"], pr["rho_hi"], pr["rho_lo"])

    def step(nthreads):
        nonlocal x
        o.set_threads(nthreads)
        o.set_spheres(x, v, w, r, ids)
        o.collide_stream(1, nthreads)
        o.sum_force_torque(x)
        x = L.advance_positions(x, v, n, wl["periodic"])

    step(threads)   # first step: builds the records
    t0 = time.perf_counter()
    for _ in range(nsteps):
        step(threads)
    dt = time.perf_counter() - t0
    t1 = time.perf_counter()
    step(1)
    dt1 = time.perf_counter() - t1
    o.close()
    cpu_model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu_model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": round(nc * nsteps / dt / 1e6, 2), "unit": "MLUPS", "cores": threads, "kind": "port",
            "single_core_value": round(nc / dt1 / 1e6, 2), "cpu_model": cpu_model,
            "sample": f"{nsteps} full coupled FP64 steps (voxelize + collide/stream + force sum, after one "
                      f"untimed step) of {wl['desc']}, by the CPU restatement of the reference algorithm on "
                      f"{threads} threads (x-slabs, as the reference's MPI ranks), {dt:.1f} s; "
                      f"single_core_value: one more step on one core, {dt1:.1f} s. The slab has the "
                      f"channel's per-cell work; MLUPS scale by cell count"}


# ---------------------------------------------------------------- profiles

PROBE_CMD = ("rocprofv3 --pmc {ctr} --output-format csv -d DIR/{tag} -o run -- "
             "python3 bench.py --kernel-probe --workload {wl}")
TRACE_CMD = ("rocprofv3 --kernel-trace --stats -d DIR/trace_{wl} -o run -- "
             "python3 bench.py --workload {wl} --only-headline --steps 10 --warmup 2 --repeats 1")
ISSUE_CTRS = "SQ_WAVES SQ_INSTS_VALU SQ_INSTS_SALU SQ_INSTS_SMEM SQ_INSTS_LDS SQ_WAVE_CYCLES SQ_BUSY_CYCLES GRBM_GUI_ACTIVE"
ISSUE_CMD = ("rocprofv3 --pmc " + ISSUE_CTRS + " --output-format csv -d DIR/issue_{wl} -o run -- "
             "python3 bench.py --kernel-probe --workload {wl}")


def _short(name):
    """k_collide_stream<PREC, FORCE, PZ> -> k_collide_stream<FP64> (+ body force / z walls noted)."""
    base = name.split("(")[0].replace("void ", "").replace("psm::", "")
    precs = {"0": "FP64", "1": "FP32_DEV", "2": "FP16_STORE"}
    if base.startswith("k_collide_stream<"):
        args = [t.strip() for t in base[len("k_collide_stream<"):].rstrip(">").split(",")]
        tag = precs.get(args[0], args[0])
        extra = []
        if len(args) > 1 and args[1] == "true":
            extra.append("body force")
        if len(args) > 2 and args[2] == "false":
            extra.append("z walls")
        return f"k_collide_stream<{tag}>" + (f" ({', '.join(extra)})" if extra else "")
    return base[:48]


def _kernel_key(sk):
    """trace/counter name -> the bench's kernel key (drops the z-walls note)."""
    return sk.split(" (")[0]


def summarize_issue(d, wl):
    import csv
    path = os.path.join(d, f"issue_{wl}", "run_counter_collection.csv")
    if not os.path.exists(path):
        return
    out = [f"# Instructions per wavefront and issue share, the coupled step's kernels ({wl})\n",
           "Counters: `" + ISSUE_CMD.format(wl=wl) + "`; one MI355X, ROCm 7.2. Counter values are",
           "summed over the device. The issue share is instructions x 4 cycles (a wave64 instruction on",
           "a 16-lane SIMD) over 1024 SIMDs x the kernel's cycles (GRBM_GUI_ACTIVE / 8, one count per XCD);",
           "VALU and SALU issue from different wavefronts in the same cycle.\n",
           "| kernel | waves | VALU / wave | SALU / wave | SMEM / wave | LDS / wave | VALU issue share | "
           "SALU issue share |", "|---|---|---|---|---|---|---|---|"]
    acc = {}
    for r in csv.DictReader(open(path)):
        acc.setdefault(r["Kernel_Name"], {}).setdefault(r["Counter_Name"], []).append(float(r["Counter_Value"]))
    rows = []
    for k, v in acc.items():
        sk = _short(k)
        if not (sk.startswith("k_collide_stream") or sk.startswith("k_fold") or sk.startswith("k_force_bin")):
            continue
        m = {c: sum(x) / len(x) for c, x in v.items()}
        waves = max(m.get("SQ_WAVES", 0.0), 1.0)
        cyc = m.get("GRBM_GUI_ACTIVE", 0.0) / 8.0
        share = (lambda n: n * 4.0 / (1024.0 * cyc)) if cyc > 0 else (lambda n: 0.0)
        rows.append((sk, waves, m.get("SQ_INSTS_VALU", 0) / waves, m.get("SQ_INSTS_SALU", 0) / waves,
                     m.get("SQ_INSTS_SMEM", 0) / waves, m.get("SQ_INSTS_LDS", 0) / waves,
                     share(m.get("SQ_INSTS_VALU", 0)), share(m.get("SQ_INSTS_SALU", 0))))
    for sk, waves, va, sa, sm, ld, vs, ss in sorted(rows):
        out.append(f"| {sk} | {waves:.0f} | {va:.0f} | {sa:.0f} | {sm:.1f} | {ld:.1f} | {vs:.2f} | {ss:.2f} |")
    with open(os.path.join(ROOT, "profiles", f"kernel_issue_counters_{wl}.md"), "w") as f:
        f.write("\n".join(out) + "\n")


def summarize_profiles(d, wl):
    """Kernel-trace stats and FETCH/WRITE counters of one workload ->
    profiles/*.md and the traffic entry of profiles/collide_stream_traffic.json."""
    import csv
    import glob
    import sqlite3
    os.makedirs(os.path.join(ROOT, "profiles"), exist_ok=True)
    dbs = glob.glob(os.path.join(d, f"trace_{wl}", "*results.db"))
    if dbs:
        con = sqlite3.connect(dbs[0])
        rows = list(con.execute("select name,total_calls,total_duration,average,percentage from top_kernels"))
        with open(os.path.join(ROOT, "profiles", f"kernels_{wl}_trace.md"), "w") as f:
            f.write(f"# Kernel durations, {WORKLOADS[wl]['desc']}\n\n")
            f.write("Command: `" + TRACE_CMD.format(wl=wl) + "` (rocprofv3 kernel trace, ROCm 7.2, one MI355X).\n")
            f.write("Per precision (FP64, FP32-dev, FP16-store) the run does 2 + 10 timed coupled steps, "
                    "then continues to the configuration's step count (200 for the channel) for the error "
                    "against FP64; after the FP64 run, the copy ceiling (k_copy16: 1 + 10 copies of the FP64 "
                    "population store). One lattice per precision: the collide's time moves by up to +-8 % "
                    "between lattices (memory placement), so the bench reports the median over fresh "
                    "lattices.\n\n")
            f.write("| kernel | calls | total us | average us | % |\n|---|---|---|---|---|\n")
            for name, n, tot, avg, pct in rows[:28]:
                f.write(f"| {_short(name)} | {n} | {tot:.1f} | {avg:.1f} | {pct:.2f} |\n")
    summarize_issue(d, wl)
    res = {}
    for tag, ctr in ((f"fetch_{wl}", "FETCH_SIZE"), (f"write_{wl}", "WRITE_SIZE")):
        path = os.path.join(d, tag, "run_counter_collection.csv")
        if not os.path.exists(path):
            continue
        acc = {}
        for r in csv.DictReader(open(path)):
            if r["Counter_Name"] != ctr:
                continue
            acc.setdefault(r["Kernel_Name"], []).append(float(r["Counter_Value"]))
        for k, v in acc.items():
            res.setdefault(k, {})[ctr] = sum(v) / len(v)
    if not res:
        return
    wlx = WORKLOADS[wl]
    nc = int(np.prod(wlx["n"]))
    # the calibration copy: k_copy16 over the FP64 population store (row-blocked, padded)
    zc_bytes = 8 * 19 * nc   # lower bound; the exact store size is printed by the probe
    probe_log = os.path.join(d, f"probe_{wl}.txt")
    try:
        zc_bytes = int(open(probe_log).read().split()[0])
    except (OSError, ValueError, IndexError):
        pass
    copy = [k for k in res if "k_copy16" in k]
    fetch_f = write_f = None
    if copy and res[copy[0]].get("FETCH_SIZE") and res[copy[0]].get("WRITE_SIZE"):
        fetch_f = zc_bytes / (res[copy[0]]["FETCH_SIZE"] * 1024.0)
        write_f = zc_bytes / (res[copy[0]]["WRITE_SIZE"] * 1024.0)
    lines = [f"# HBM traffic per dispatch, {wlx['desc']}\n",
             "Counters: `" + PROBE_CMD.format(ctr="FETCH_SIZE", tag=f"fetch_{wl}", wl=wl) +
             "`, then the same with WRITE_SIZE (tag write_" + wl + ").",
             f"Calibration: the copy k_copy16 reads and writes the FP64 population store, {zc_bytes} B each way;",
             (f"FETCH_SIZE scale = {fetch_f:.4f} B per counted KiB/1024, WRITE_SIZE scale = {write_f:.4f}.\n"
              if fetch_f else "no calibration copy found: unscaled.\n"),
             "| kernel | FETCH_SIZE KiB | WRITE_SIZE KiB | read GB (scaled) | written GB (scaled) | algorithmic GB | measured / algorithmic |",
             "|---|---|---|---|---|---|---|"]
    traffic = {}
    algo = {KERNEL[p]: 2 * 19 * BYTES_PER_POP[p] * nc for p in PRECISIONS}
    algo["k_copy16"] = 2 * zc_bytes
    for k, v in sorted(res.items(), key=lambda kv: -kv[1].get("FETCH_SIZE", 0)):
        fk, wk = v.get("FETCH_SIZE", 0.0), v.get("WRITE_SIZE", 0.0)
        rd = fk * 1024 * (fetch_f or 1) / 1e9
        wr = wk * 1024 * (write_f or 1) / 1e9
        sk = _short(k)
        al = algo.get(_kernel_key(sk))
        if al:
            lines.append(f"| {sk} | {fk:.0f} | {wk:.0f} | {rd:.3f} | {wr:.3f} | {al / 1e9:.3f} | {(rd + wr) * 1e9 / al:.3f} |")
        else:
            lines.append(f"| {sk} | {fk:.0f} | {wk:.0f} | {rd:.3f} | {wr:.3f} | | |")
        if sk.startswith("k_collide_stream") or sk.startswith("k_copy16"):
            traffic[_kernel_key(sk)] = (rd + wr) * 1e9
    with open(os.path.join(ROOT, "profiles", f"collide_stream_{wl}_hbm_traffic.md"), "w") as f:
        f.write("\n".join(lines[:40]) + "\n")
    try:
        t = json.load(open(TRAFFIC_JSON))
    except (OSError, ValueError):
        t = {}
    t.setdefault("workloads", {})[wl] = {"bytes_per_launch": traffic, "fetch_scale": fetch_f,
                                         "write_scale": write_f, "source": f"collide_stream_{wl}_hbm_traffic.md"}
    for stale in ("workload", "bytes_per_launch", "fetch_scale", "write_scale"):
        t.pop(stale, None)
    with open(TRAFFIC_JSON, "w") as f:
        json.dump(t, f, indent=1)


ADJ_TRACE_CMD = ("rocprofv3 --kernel-trace --stats -d DIR/trace_adjoint128 -o run -- "
                 "python3 bench.py --adjoint-only")


def summarize_adjoint_trace(d):
    """Kernel-trace stats of the --adjoint-only run -> profiles/kernels_adjoint128_trace.md."""
    import glob
    import sqlite3
    dbs = glob.glob(os.path.join(d, "trace_adjoint128", "*results.db"))
    if not dbs:
        return
    con = sqlite3.connect(dbs[0])
    rows = list(con.execute("select name,total_calls,total_duration,average,percentage from top_kernels"))
    with open(os.path.join(ROOT, "profiles", "kernels_adjoint128_trace.md"), "w") as f:
        f.write(f"# Kernel durations, {ADJOINT_CASE['desc']}\n\n")
        f.write("Command: `" + ADJ_TRACE_CMD + "` (rocprofv3 kernel trace, ROCm 7.2, one MI355X).\n")
        f.write("Per precision (FP64, then FP32-dev) the run does one window to allocate the state pool "
                "(500 steps forward with a checkpoint every 50, then the backward pass), the timed window "
                "of the same shape, and 2 + 500 plain steps of the same frozen geometry. Every collide "
                "here runs with the body force.\n\n")
        f.write("| kernel | calls | total us | average us | % |\n|---|---|---|---|---|\n")
        for name, n, tot, avg, pct in rows[:24]:
            f.write(f"| {_short(name)} | {n} | {tot:.1f} | {avg:.1f} | {pct:.2f} |\n")


ADJ_C5_TRACE_CMD = ("rocprofv3 --kernel-trace --stats -d DIR/trace_adjoint_channel512 -o run -- "
                    "python3 bench.py --adjoint-window-only")
ADJ_C5_PMC_CMD = ("rocprofv3 --pmc {ctr} --output-format csv -d DIR/{tag}_adjoint_channel512 -o run -- "
                  "python3 bench.py --adjoint-window-only --probe-out DIR")


def summarize_adjoint_c5(d):
    """Kernel trace and HBM counters of --adjoint-window-only ->
    profiles/kernels_adjoint_channel512_trace.md: per phase (the marker
    copies k_copy16 split forward, backward and plain steps) the kernels'
    time, and bytes per dispatch of the backward's kernels."""
    import csv
    import glob
    import sqlite3
    dbs = glob.glob(os.path.join(d, "trace_adjoint_channel512", "*results.db"))
    if not dbs:
        return
    con = sqlite3.connect(dbs[0])
    rows = list(con.execute("select name, start, end from kernels order by start"))
    marks = [i for i, r in enumerate(rows) if "k_copy16" in r[0]]
    if len(marks) < 6:
        return
    # each marker is copy_bandwidth(1): a warm-up copy and a timed one
    phases = (("forward (50 steps, budget 10)", marks[1] + 1, marks[2]),
              ("backward", marks[3] + 1, marks[4]),
              ("plain frozen steps (50)", marks[5] + 1, len(rows)))
    c = ADJOINT_C5
    lines = [f"# Kernel durations, {c['desc'].split(', d(total')[0]}: the 50-step window\n",
             "Command: `" + ADJ_C5_TRACE_CMD + "` (rocprofv3 kernel trace, ROCm 7.2, one MI355X); counters: `" +
             ADJ_C5_PMC_CMD.format(ctr="FETCH_SIZE", tag="fetch") + "`, then WRITE_SIZE (tag write).",
             "The run: a warm-up window (state pool, adjoint buffers), 150 coupled steps, then the window's "
             "forward, its backward (d(total z-drag)) and 50 plain steps of the same frozen geometry, with a "
             "copy of the population store (k_copy16) between the phases to mark them.\n"]
    for label, a, b in phases:
        acc = {}
        for name, s, e in rows[a:b]:
            k = _short(name)
            v = acc.setdefault(k, [0, 0.0])
            v[0] += 1
            v[1] += (e - s) / 1e6
        span = (rows[b - 1][2] - rows[a][1]) / 1e6
        tot = sum(v[1] for v in acc.values())
        lines += [f"## {label}: {span:.1f} ms from first to last kernel, kernels {tot:.1f} ms\n",
                  "| kernel | calls | total ms | average ms |", "|---|---|---|---|"]
        for k, v in sorted(acc.items(), key=lambda kv: -kv[1][1])[:10]:
            lines.append(f"| {k} | {v[0]} | {v[1]:.2f} | {v[1] / v[0]:.3f} |")
        lines.append("")
    res = {}
    for tag, ctr in (("fetch", "FETCH_SIZE"), ("write", "WRITE_SIZE")):
        path = os.path.join(d, f"{tag}_adjoint_channel512", "run_counter_collection.csv")
        if not os.path.exists(path):
            continue
        for r in csv.DictReader(open(path)):
            if r["Counter_Name"] == ctr:
                res.setdefault(_short(r["Kernel_Name"]), {}).setdefault(ctr, []).append(float(r["Counter_Value"]))
    try:
        store = int(open(os.path.join(d, "probe_adjoint_channel512.txt")).read().split()[0])
    except (OSError, ValueError, IndexError):
        store = None
    cp = res.get("k_copy16", {})
    if store and cp.get("FETCH_SIZE") and cp.get("WRITE_SIZE"):
        ff = store / (np.mean(cp["FETCH_SIZE"]) * 1024.0)
        wf = store / (np.mean(cp["WRITE_SIZE"]) * 1024.0)
        lines += [f"## HBM bytes per dispatch (scaled by the marker copy of the FP32 store, {store} B each way: "
                  f"FETCH_SIZE x {ff:.4f}, WRITE_SIZE x {wf:.4f})\n",
                  "| kernel | dispatches | read GB | written GB |", "|---|---|---|---|"]
        for k, v in sorted(res.items(), key=lambda kv: -np.mean(kv[1].get("FETCH_SIZE", [0]))):
            if "FETCH_SIZE" not in v or np.mean(v["FETCH_SIZE"]) * 1024 * ff < 1e8:
                continue
            lines.append(f"| {k} | {len(v['FETCH_SIZE'])} | {np.mean(v['FETCH_SIZE']) * 1024 * ff / 1e9:.2f} | "
                         f"{np.mean(v.get('WRITE_SIZE', [0])) * 1024 * wf / 1e9:.2f} |")
    with open(os.path.join(ROOT, "profiles", "kernels_adjoint_channel512_trace.md"), "w") as f:
        f.write("\n".join(lines) + "\n")


def kernel_probe(key, out_dir=None):
    """Short run for hardware-counter profiles: the copy k_copy16 of the FP64
    population store (known size, for calibration), then 3 coupled steps per
    precision."""
    wl = WORKLOADS[key]
    sph = make_spheres(wl)
    lat = new_lattice(wl, "fp64")
    ms, nbytes = lat.copy_bandwidth(1)
    lat.close()
    log(f"probe copy: {int(nbytes // 2)} bytes read and written per copy, {ms:.3f} ms")
    if out_dir:
        os.makedirs(out_dir, exist_ok=True)
        with open(os.path.join(out_dir, f"probe_{key}.txt"), "w") as f:
            f.write(f"{int(nbytes // 2)}\n")
    for prec in PRECISIONS:
        run_gpu(wl, prec, 3, 1, sph)
        log(f"probe {prec} done")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # the measurement protocol: >= 20 warm-up and >= 100 timed steps, median of 5 fresh lattices
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--repeats", type=int, default=5,
                    help="fresh lattices per precision of the headline workload; value is the median")
    ap.add_argument("--workload", default="channel512", choices=sorted(k for k in WORKLOADS if k != "channel512_slab"))
    ap.add_argument("--only-headline", action="store_true",
                    help="the headline workload only (no other workloads, adjoint, exchange loop or CPU baseline)")
    ap.add_argument("--no-extra", action="store_true", help="skip the other workloads")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the differentiable-window lines")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="only the 128x128x256 differentiable window, FP64 and FP32-dev (for its kernel trace); "
                         "with --summarize-profiles DIR: summarize DIR/trace_adjoint128")
    ap.add_argument("--adjoint-window-only", action="store_true",
                    help="only the timed 512x512x768 adjoint window (forward, backward, plain steps; no central "
                         "differences), with a marker copy between the phases (for its kernel trace and "
                         "counters); with --summarize-profiles DIR: summarize DIR/trace_adjoint_channel512, "
                         "DIR/fetch_adjoint_channel512 and DIR/write_adjoint_channel512")
    ap.add_argument("--cpu-steps", type=int, default=6)
    ap.add_argument("--summarize-profiles", metavar="DIR",
                    help="turn rocprofv3 outputs under DIR (trace_<wl>/, fetch_<wl>/, write_<wl>/, "
                         "issue_<wl>/) into profiles/*.md and profiles/collide_stream_traffic.json")
    ap.add_argument("--kernel-probe", action="store_true",
                    help="short run for hardware-counter profiles: the calibration copy of known size, "
                         "then 3 coupled steps per precision")
    ap.add_argument("--probe-out", default=os.path.join(ROOT, "gpurun_out"),
                    help="where --kernel-probe writes the calibration copy's size")
    args = ap.parse_args()
    if args.summarize_profiles:
        if args.adjoint_window_only:
            summarize_adjoint_c5(args.summarize_profiles)
        elif args.adjoint_only:
            summarize_adjoint_trace(args.summarize_profiles)
        else:
            summarize_profiles(args.summarize_profiles, args.workload)
        return
    if args.gpus != 1:
        print(f"bench.py: the library runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU slab decomposition is a design, not a measured path)", file=sys.stderr)
        sys.exit(2)
    if args.kernel_probe:
        kernel_probe(args.workload, args.probe_out)
        return
    if args.adjoint_window_only:
        wl, c = WORKLOADS["channel512"], ADJOINT_C5
        sph = make_spheres(wl)
        xw = np.array(sph[0], dtype=np.float64)
        for _ in range(c["pre_steps"]):
            xw = L.advance_positions(xw, sph[1], wl["n"], wl["periodic"])
        fwd, bwd, plain, _, _, mem, state_gb = time_adjoint_c5(sph, xw, markers=True)
        os.makedirs(args.probe_out, exist_ok=True)
        with open(os.path.join(args.probe_out, "probe_adjoint_channel512.txt"), "w") as f:
            f.write(f"{int(round(state_gb * 1e9 / 4)) * 4}\n")
        print(json.dumps({"workload": c["desc"], "forward_s": round(fwd, 4), "backward_s": round(bwd, 4),
                          "plain_forward_s": round(plain, 4), "device_memory_used_after_backward_gb": mem[0] if mem else None}))
        return
    if args.adjoint_only:
        adj = {}
        for prec in ("fp64", "fp32_dev"):
            log(f"adjoint window {prec}: {ADJOINT_CASE['steps']} steps forward + backward")
            adj[prec] = run_adjoint(prec)
            adj[prec].pop("_g")
            log(f"adjoint {prec}: forward {adj[prec]['forward_s']:.3f} s, backward {adj[prec]['backward_s']:.3f} s")
        print(json.dumps({"workload": ADJOINT_CASE["desc"], **adj}))
        return

    key = args.workload
    head, res, sph = run_workload(key, args.steps, args.warmup, args.repeats, copy_reps=10)
    f64 = res["fp64"]
    extra_on = not (args.only_headline or args.no_extra)
    adj_on = not (args.only_headline or args.no_adjoint)
    extra = {}
    if adj_on and key == "channel512":
        log("channel512: adjoint of the last 50 of 200 steps (FP32, state budget 10)")
        try:
            head["adjoint"] = run_adjoint_c5(sph)
            a5 = head["adjoint"]
            log(f"channel512 adjoint: forward {a5['forward_s']:.3f} s, backward {a5['backward_s']:.3f} s, "
                f"device memory {a5['device_memory_used_after_backward_gb']} GB")
        except L.PSMError as e:
            head["adjoint"] = {"error": str(e)}
            log(f"channel512 adjoint failed: {e}")
    if not args.only_headline:
        ex = {}
        for prec in ("fp64", "fp32_dev"):
            log(f"{key} exchange loop {prec}")
            ex[prec] = run_exchange(key, prec, max(5, args.steps // 5), 2, sph)
            ex[prec]["resident_step_mlups"] = head[prec]["mlups"]
            log(f"  {ex[prec]['mlups']:.0f} MLUPS")
        head["exchange_loop"] = ex
    del res
    if extra_on:
        # the channel's slab first: one GPU's share of an 8-way split along x
        for k2 in ("channel512_slab", "cube512", "dense256", "poly384"):
            if k2 == key:
                continue
            steps2 = args.steps
            w2 = WORKLOADS[k2]
            if k2 == "poly384":
                # its configuration's 1000 steps: 2 warm-up, the rest timed
                extra[k2], _, _ = run_workload(k2, w2["full_steps"] - 2, 2, 1)
            else:
                extra[k2], _, _ = run_workload(k2, steps2, args.warmup, 1)
    if adj_on:
        adj = {}
        for prec in ("fp64", "fp32_dev"):
            log(f"adjoint window {prec}: {ADJOINT_CASE['steps']} steps forward + backward")
            adj[prec] = run_adjoint(prec)
            log(f"adjoint {prec}: forward {adj[prec]['forward_s']:.3f} s, backward {adj[prec]['backward_s']:.3f} s")
        g64, g32 = adj["fp64"].pop("_g"), adj["fp32_dev"].pop("_g")
        extra["adjoint"] = {"workload": ADJOINT_CASE["desc"], "fp64": adj["fp64"], "fp32_dev": adj["fp32_dev"],
                            "fp32_grad_rel_err_vs_fp64": {k: rel(g32[k], g64[k]) for k in ("v", "w", "x", "sf")}}

    cpu = None
    if args.cpu_steps > 0 and not args.only_headline:
        log(f"cpu baseline: {args.cpu_steps} oracle steps on the channel slab")
        cpu = cpu_baseline(args.cpu_steps)
        log(f"cpu baseline: {cpu['value']:.1f} MLUPS on {cpu['cores']} threads")

    wl = WORKLOADS[key]
    nc = int(np.prod(wl["n"]))
    h64 = head["fp64"]
    out = {
        "metric": METRIC,
        "value": h64["mlups"],
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "repeats": max(1, args.repeats),
        "value_is": "median over fresh lattices of the FP64 full-step rate (each: warmup, then exactly "
                    "`steps` timed steps)",
        "ms_per_step": h64["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": wl["desc"], "cells": nc, "spheres": wl["count"], "sub_cell_samples": 8,
                   "kinematics": "prescribed x += v", "precision": "fp64",
                   "step": "voxelize + fused collide/stream + force/torque sum"},
        "roofline": dict(h64["roofline"], device_copy_gbs=h64.get("device_copy_gbs"),
                         device_copy="k_copy16 over the FP64 population store of the same lattice, "
                                     "read + write bytes per second (HIP events)",
                         traffic_source=f"rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE scaled by the k_copy16 "
                                        f"copy of known size; profiles/collide_stream_{key}_hbm_traffic.md"),
        "cpu_baseline": cpu,
        "fp64": h64,
        "fp32_dev": head["fp32_dev"],
        "fp16_store": head["fp16_store"],
        "fp32_over_fp64_mlups": head["fp32_over_fp64_mlups"],
        "fp16_over_fp64_mlups": head["fp16_over_fp64_mlups"],
    }
    for k in ("adjoint", "exchange_loop"):
        if k in head:
            out[f"{key}_{k}"] = head[k]
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
