"""The reference's flow configurations at full size, driven through the C ABI.

Each builder returns the LUMA-layout set-up (labels, inlet, initial state, refined region,
markers) as LUMA's initialisation would produce it; each runner steps it with
``LBM_multi_opt`` and reports the quantities each configuration is measured by (fields, drag, lift,
Strouhal number, throughput).  Command line::

    python -m luma_amd.cases c1 --precision fp64 --steps 20000
    python -m luma_amd.cases c3 --precision fp64,fp32s --steps 100000 --out c3.json

Lattice units throughout.  Drag and lift coefficients use the mean inlet velocity
U0 = Umax / 1.5 (LUMA's L_UX0, on which L_RE is based) and the diameter in cells of the level
carrying the body:  C = 2 F / (rho0 U0^2 D).  The body force is the FP64 sum LUMA writes to
Body_<id>_LD (ObjectManager_ops_io.cpp:98-121).
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np

from . import setups
from .grid import GpuGrid


# ---- C1: D2Q9 lid-driven cavity 256^2, tau 0.56, lid 0.1, HWBB walls --------------------

def c1_cavity(n: int = 256) -> setups.Case:
    """L_RESOLUTION 256, L_RE 1280, lid u = 0.1: tau = 0.56 (GridObj_init_grids.cpp:339-344)."""
    return setups.lid_driven_cavity(n, re=1280.0 * n / 256, u_lid_lbm=0.1)


def run_c1(precision: str, steps: int, n: int = 256):
    case = c1_cavity(n)
    with GpuGrid.from_case(case, precision, store_macro=True) as g:
        t0 = time.perf_counter()
        g.run(steps)
        g.sync()
        wall = time.perf_counter() - t0
        rho, u = g.download_macro()
    return {"config": "c1_cavity", "precision": precision, "steps": steps, "N": n,
            "wall_s": wall, "mlups": case.N * case.M * steps / wall / 1e6}, (rho, u, case)


# ---- C3: D2Q9 channel, IB cylinder Re 100, one refined level -----------------------------

C3_RES = 40.0                     # coarse cells per cylinder diameter (LUMA's unit length)
C3_UMAX = 0.04                    # parabolic inlet maximum, lattice units
C3_RE = 100.0                     # on the mean inlet velocity and the diameter
C3_CENTRE = (256.0, 130.0)        # coarse cells: x = 256, y = 128 + 2 (the seeded offset)
C3_PATCH = (208, 400, 90, 170)    # coarse cells covered by the fine level
C3_MARKERS = 252


def c3_cylinder(nx: int = 1024, ny: int = 256, patch=C3_PATCH, centre=C3_CENTRE, nmark: int = C3_MARKERS):
    """Coarse 1024 x 256 channel (parabolic inlet, ExtrapolateRight outlet, solid walls) with a
    2:1 fine patch around a 40-cell cylinder of ``nmark`` equally spaced markers on the fine level."""
    parent = setups.channel_2d(nx, ny, C3_RES, C3_UMAX, C3_RE)
    dh, dt = parent.info["dh"], parent.info["dt"]
    child = setups.add_refined_region(parent, patch, dh, dt)
    setups.velocity_field_on(child, parent)
    pos = setups.circle_markers((centre[0] * dh, centre[1] * dh, 0.0), 0.5 * C3_RES * dh, child.info["dh"],
                                luma_skip_first=False, n=nmark)
    return parent, child, pos


def strouhal(cl: np.ndarray, dstep: float, diameter: float, u0: float, tail: float = 0.5):
    """St = f D / U0 with f the dominant frequency of the lift signal over its last ``tail``
    fraction (FFT peak of the mean-removed, Hann-windowed signal, refined by a parabola
    through the peak and its two neighbours).  Returns (St, period in steps), or (None, None)
    for a window too short or flat to the last bits (a steady wake)."""
    x = cl[int(len(cl) * (1.0 - tail)):]
    if len(x) < 16 or np.ptp(x) <= 1e-12 * max(1.0, float(np.abs(x).max())):
        return None, None
    x = (x - x.mean()) * np.hanning(len(x))
    p = np.abs(np.fft.rfft(x)) ** 2
    k = int(np.argmax(p[1:])) + 1
    if 1 <= k < len(p) - 1:
        a, b, c = np.log(p[k - 1] + 1e-300), np.log(p[k]), np.log(p[k + 1] + 1e-300)
        k = k + 0.5 * (a - c) / (a - 2 * b + c)
    period = len(x) * dstep / k
    return diameter / (period * u0), period


def run_c3(precision: str, steps: int, sample: int = 20, fields_at=(), geometry=None):
    parent, child, pos = geometry if geometry is not None else c3_cylinder()
    u0 = C3_UMAX / 1.5
    d_fine = C3_RES * 2.0
    snaps = {}
    cd, cl = [], []
    with GpuGrid.from_case(parent, precision, store_macro=True) as gp, \
            GpuGrid.from_case(child, precision, store_macro=True) as gc:
        gp.add_child(gc)
        body = gc.ibm_add_body(pos)
        gc.ibm_initialise()
        t0 = time.perf_counter()
        for s in range(sample, steps + 1, sample):
            gp.run(sample)
            F = gc.ibm_body_force(body)
            cd.append(2.0 * F[0] / (u0 * u0 * d_fine))
            cl.append(2.0 * F[1] / (u0 * u0 * d_fine))
            if s in fields_at:
                snaps[s] = (gp.download_macro(), gc.download_macro())
            if s % 10000 == 0:
                print(f"[c3 {precision}] step {s} Cd {cd[-1]:.5f} Cl {cl[-1]:+.5f}", flush=True)
        gp.sync()
        wall = time.perf_counter() - t0
    cd, cl = np.array(cd), np.array(cl)
    st, period = strouhal(cl, sample, C3_RES, u0)
    half = cd[len(cd) // 2:]
    cells = parent.N * parent.M + 2 * child.N * child.M
    out = {"config": "c3_cylinder", "precision": precision, "steps": steps, "markers": len(pos),
           "fine_patch": [child.N, child.M], "tau_coarse": 1.0 / parent.omega, "tau_fine": 1.0 / child.omega,
           "Cd_mean": float(half.mean()), "Cd_amp": float((half.max() - half.min()) / 2),
           "Cl_amp": float((cl[len(cl) // 2:].max() - cl[len(cl) // 2:].min()) / 2),
           "St": st, "period_steps": period, "wall_s": wall, "mlups": cells * steps / wall / 1e6}
    return out, {"cd": cd, "cl": cl, "snaps": snaps, "parent": parent, "child": child}


# ---- C4: D3Q19 uniform inflow past an IB sphere, adjoint of the time-averaged drag --------

C4_U = 0.03
C4_RE = 50.0                      # on U and the sphere diameter
C4_DIAM = 24.0
C4_CENTRE = (96.0, 64.0, 64.0)
C4_MARKERS = 1810
C4_SEED = 20261202
C4_SIGMA = 0.05                   # marker position noise, cells


def c4_markers(seed: int = C4_SEED, n: int = C4_MARKERS, centre=C4_CENTRE, diam: float = C4_DIAM):
    """Fibonacci sphere of n markers (Body.h:215-228 with n given) with seeded N(0, 0.05 cell)
    noise on every coordinate (numpy PCG64)."""
    base = setups.sphere_markers(centre, diam / 2.0, 1.0, n=n)
    return base + np.random.default_rng(seed).normal(0.0, C4_SIGMA, base.shape)


def c4_tau(U: float = C4_U) -> float:
    return 3.0 * U * C4_DIAM / C4_RE + 0.5


def c4_case(U: float = C4_U, tau: float = None, shape=(384, 128, 128)):
    return setups.uniform_inflow_3d(*shape, U, tau if tau is not None else c4_tau(C4_U))


def c4_eval(precision, U, tau, pos, steps, every, backward, shape=(384, 128, 128)):
    case = c4_case(U, tau, shape)
    with GpuGrid.from_case(case, precision) as g:
        g.ibm_add_body(pos)
        g.ibm_initialise()
        t0 = time.perf_counter()
        out = g.adj_run(steps, every, backward=backward, positions=backward, initial=backward)
        out["wall_s"] = time.perf_counter() - t0
    if backward:
        # U enters through the inlet (scale U/U0) and the uniform initial state feq(1, U e_x)
        c = np.array([setups.lattice_c(19, v) for v in range(19)], dtype=float)
        w = np.array([setups.lattice_w(19, v) for v in range(19)])
        dfeq = w * (3.0 * c[:, 0] + 9.0 * (c[:, 0] * U) * c[:, 0] - 3.0 * U)
        out["dJ_dU"] = out["dJ_dinlet"] / U + float((out.pop("lambda0") * dfeq).sum())
        out["dJ_dtau"] = -out["dJ_domega"] / (tau * tau)
    return out


def run_c4_forward(precision: str, steps: int, body: bool = True, warmup: int = 20):
    """Forward-step throughput of the C4 arrangement (384 x 128 x 128, forced-equilibrium inlet,
    ExtrapolateRight outlet, the 1,810-marker sphere when ``body``): ``steps`` LBM_multi_opt calls
    after ``warmup``, between two synchronisations."""
    case = c4_case()
    with GpuGrid.from_case(case, precision) as g:
        if body:
            g.ibm_add_body(c4_markers())
            g.ibm_initialise()
        g.run(warmup)
        g.sync()
        t0 = time.perf_counter()
        g.run(steps)
        g.sync()
        wall = time.perf_counter() - t0
    cells = case.N * case.M * case.K
    return {"config": "c4_forward", "precision": precision, "body": body, "steps": steps, "wall_s": wall,
            "ms_per_step": 1e3 * wall / steps, "mlups": cells * steps / wall / 1e6}


def run_c4_adjoint_timing(precision: str, steps: int, every: int = 200):
    """Wall clock of the C4 gradient alone: ``steps`` forward steps with checkpoints, then the
    discrete adjoint of them (segments recomputed from their checkpoints), no differences."""
    pos = c4_markers()
    a = c4_eval(precision, C4_U, c4_tau(), pos, steps, every, True)
    cells = 384 * 128 * 128
    return {"config": "c4_adjoint_timing", "precision": precision, "steps": steps, "checkpoint_every": every,
            "J": a["J"], "dJ_dU": a["dJ_dU"], "dJ_dtau": a["dJ_dtau"], "wall_s": a["wall_s"],
            "ms_per_step": 1e3 * a["wall_s"] / steps, "mlups_forward_plus_adjoint": cells * steps / a["wall_s"] / 1e6}


def run_c4(precisions=("fp64", "fp32s"), steps=4000, every=200, h=1e-4, fd_markers=(0, 605, 1210),
           shape=(384, 128, 128)):
    pos = c4_markers()
    U, tau = C4_U, c4_tau()
    rep = {"config": "c4_sphere_adjoint", "shape": list(shape), "markers": len(pos), "U": U, "tau": tau,
           "steps": steps, "checkpoint_every": every, "fd_h": h}
    for p in precisions:
        a = c4_eval(p, U, tau, pos, steps, every, True, shape)
        rep[p] = {"J": a["J"], "dJ_dU": a["dJ_dU"], "dJ_dtau": a["dJ_dtau"], "wall_s": a["wall_s"],
                  "dJ_dpos_fd_markers": a["dJ_dpos"][list(fd_markers)].tolist(),
                  "dJ_dpos_norm": float(np.linalg.norm(a["dJ_dpos"]))}
        print(f"[c4 {p}] J {a['J']:.9e} dJ/dU {a['dJ_dU']:.9e} dJ/dtau {a['dJ_dtau']:.9e} ({a['wall_s']:.1f} s)",
              flush=True)
    # central differences of the FP64 forward
    J = lambda U_, tau_, pos_: c4_eval("fp64", U_, tau_, pos_, steps, every, False, shape)["J"]
    fd = {"dJ_dU": (J(U * (1 + h), tau, pos) - J(U * (1 - h), tau, pos)) / (2 * h * U),
          "dJ_dtau": (J(U, tau * (1 + h), pos) - J(U, tau * (1 - h), pos)) / (2 * h * tau)}
    print(f"[c4 fd] dJ/dU {fd['dJ_dU']:.9e} dJ/dtau {fd['dJ_dtau']:.9e}", flush=True)
    pfd = []
    for m in fd_markers:
        row = []
        for a_ in range(3):
            pp, pm = pos.copy(), pos.copy()
            pp[m, a_] += h
            pm[m, a_] -= h
            row.append((J(U, tau, pp) - J(U, tau, pm)) / (2 * h))
        pfd.append(row)
        print(f"[c4 fd] dJ/dx[{m}] {row}", flush=True)
    fd["dJ_dpos_fd_markers"] = pfd
    rep["fd_fp64"] = fd
    ref = rep["fp64"] if "fp64" in rep else None
    for p in precisions:
        r = rep[p]
        gp = np.array(r["dJ_dpos_fd_markers"])
        scale = np.abs(np.array(pfd)).max()
        r["rel_err_vs_fd"] = {"dJ_dU": abs(r["dJ_dU"] - fd["dJ_dU"]) / abs(fd["dJ_dU"]),
                              "dJ_dtau": abs(r["dJ_dtau"] - fd["dJ_dtau"]) / abs(fd["dJ_dtau"]),
                              "dJ_dpos_max_over_max": float(np.abs(gp - np.array(pfd)).max() / scale)}
        if ref is not None and p != "fp64":
            r["rel_err_vs_fp64_adjoint"] = {
                "dJ_dU": abs(r["dJ_dU"] - ref["dJ_dU"]) / abs(ref["dJ_dU"]),
                "dJ_dtau": abs(r["dJ_dtau"] - ref["dJ_dtau"]) / abs(ref["dJ_dtau"])}
    return rep


# ---- C5: FEM-coupled flexible filament in a D2Q9 channel with a refined patch -------------

C5_RES = 100.0                    # coarse cells per filament length (LUMA's unit length)
C5_UMAX = 0.05
C5_RE = 200.0                     # on the mean inlet velocity and the filament length
C5_PATCH = (450, 800, 180, 331)   # coarse cells covered by the fine level
C5_LEAD = (500.0, 256.0)          # clamped leading end, coarse cells
C5_HEIGHT = 0.01                  # filament thickness, units of length
C5_ELEMENTS = 40
C5_SEED = 20261202


def c5_tip_deflection(seed: int = C5_SEED) -> float:
    """Seeded initial tip deflection, U(-2, 2) fine cells."""
    return float(np.random.default_rng(seed).uniform(-2.0, 2.0))


def c5_filament(nx: int = 2048, ny: int = 512, patch=C5_PATCH):
    """Coarse 2048 x 512 channel (parabolic inlet, outlet, solid walls) with a 2:1 fine patch
    around a clamped filament of 200 fine cells (201 markers, 40 corotational elements),
    density ratio 10, dimensionless bending stiffness 1e-3.  LUMA starts the filament straight
    (FEMBody.cpp:95-100); the seeded initial tip deflection is given here as the filament's
    initial angle."""
    parent = setups.channel_2d(nx, ny, C5_RES, C5_UMAX, C5_RE)
    dh, dt = parent.info["dh"], parent.info["dt"]
    child = setups.add_refined_region(parent, patch, dh, dt)
    setups.velocity_field_on(child, parent)
    dhf = child.info["dh"]
    length = 1.0
    tip = c5_tip_deflection()
    angle = math.degrees(math.asin(tip * dhf / length))
    density, E = setups.filament_material(length, C5_HEIGHT, dhf, mass_ratio=10.0, bending=1e-3)
    fil = {"start": (C5_LEAD[0] * dh, C5_LEAD[1] * dh + 0.0013, 0.0), "length": length, "height": C5_HEIGHT,
           "angle": angle, "elements": C5_ELEMENTS, "density": density, "E": E, "tip_cells": tip}
    return parent, child, fil


def run_c5(precision: str, steps: int, sample: int = 1):
    parent, child, fil = c5_filament()
    tips, its = [], []
    with GpuGrid.from_case(parent, precision) as gp, GpuGrid.from_case(child, precision) as gc:
        gp.add_child(gc)
        b = gc.ibm_add_filament(fil["start"], fil["length"], fil["height"], fil["angle"], fil["elements"], True,
                                fil["density"], fil["E"])
        gc.ibm_initialise()
        t0 = time.perf_counter()
        for s in range(1, steps + 1):
            gp.LBM_multi_opt()
            its.append(gc.ibm_subiterations()[0])
            if s % sample == 0:
                tips.append(gc.ibm_markers(b)["pos"][-1, 1])
            if s % 2000 == 0:
                print(f"[c5 {precision}] coarse step {s} tip y {tips[-1]:.6f} sub-iterations {its[-1]}", flush=True)
        gp.sync()
        wall = time.perf_counter() - t0
    tips = np.array(tips)
    dhf = child.info["dh"]
    y0 = fil["start"][1]
    tail = (tips[len(tips) // 2:] - y0) / dhf
    cells = parent.N * parent.M + 2 * child.N * child.M
    return {"config": "c5_filament", "precision": precision, "coarse_steps": steps, "fine_steps": 2 * steps,
            "markers": int(np.floor(fil["length"] / dhf)) + 1, "elements": fil["elements"],
            "initial_tip_cells": fil["tip_cells"], "tau_coarse": 1.0 / parent.omega, "tau_fine": 1.0 / child.omega,
            "tip_amplitude_cells": float((tail.max() - tail.min()) / 2), "tip_mean_cells": float(tail.mean()),
            "mean_subiterations": float(np.mean(its)), "wall_s": wall, "ms_per_coarse_step": 1e3 * wall / steps,
            "mlups": cells * steps / wall / 1e6}, tips


def c5_tip_series(precision: str, steps: int, youngs: float, seed: bool = False, pin=None):
    """C5 tip trajectory: the last marker's y per coarse step, its derivative along Young's
    modulus (tangent form, seeded), and the FSI sub-iteration count of every fine step (pin:
    a schedule of counts to follow instead of the residual test)."""
    parent, child, fil = c5_filament()
    tips, dtips, its = np.empty(steps), np.empty(steps), []
    with GpuGrid.from_case(parent, precision) as gp, GpuGrid.from_case(child, precision) as gc:
        gp.add_child(gc)
        b = gc.ibm_add_filament(fil["start"], fil["length"], fil["height"], fil["angle"], fil["elements"], True,
                                fil["density"], youngs)
        gc.ibm_initialise()
        if seed:
            gc.ibm_seed_youngs(b, 1.0)
        if pin is not None:
            gc.ibm_pin_subiterations(pin)
        t0 = time.perf_counter()
        for s in range(steps):
            gp.LBM_multi_opt()
            tips[s] = gc.ibm_markers(b)["pos"][-1, 1]
            if seed:
                dtips[s] = gc.ibm_marker_tangents(b)["pos"][-1, 1]
            if (s + 1) % 2000 == 0:
                print(f"[c5 grad {precision}] coarse step {s + 1} tip y {tips[s]:.6f}", flush=True)
        wall = time.perf_counter() - t0
        its = gc.ibm_subiteration_log()
    return tips, (dtips if seed else None), its, wall


def c5_objectives(tips, dtips, y0, dhf):
    """Mean tip amplitude J = mean_t |y_tip - y0| / dh_f over the second half of the run (the
    window run_c5 reports), the mean signed deflection, and their derivatives when dtips is
    given (per unit of the seeded parameter)."""
    w = slice(len(tips) // 2, None)
    dev = (tips[w] - y0) / dhf
    out = {"J_mean_amplitude_cells": float(np.abs(dev).mean()), "J_mean_deflection_cells": float(dev.mean())}
    if dtips is not None:
        ddev = dtips[w] / dhf
        out["dJ_mean_amplitude"] = float((np.sign(dev) * ddev).mean())
        out["dJ_mean_deflection"] = float(ddev.mean())
    return out


def run_c5_grad(steps: int, check_steps: int, h_rel: float = 1e-3, form: str = "fp64t"):
    """d(mean tip amplitude)/d(bending stiffness) for C5 in the tangent form, with the tip
    trajectory's derivative checked against central differences of FP64 runs at E(1 +- h_rel)
    over the first check_steps coarse steps.  The FSI loop stops on a residual threshold, so
    the difference runs follow the derivative run's sub-iteration schedule (pinned): both
    then differentiate the same discrete program."""
    parent, child, fil = c5_filament()
    E, KB = fil["E"], 1e-3
    y0, dhf = fil["start"][1], child.info["dh"]
    tips, dtips, its, wall = c5_tip_series(form, steps, E, seed=True)
    obj = c5_objectives(tips, dtips, y0, dhf)
    # E = KB * rho_f u^2 L^3 / I (setups.filament_material): dJ/dKB = dJ/dE * E / KB
    rep = {"config": "c5_filament_gradient", "form": form,
           "coarse_steps": steps, "fine_steps": 2 * steps, "bending_stiffness": KB, "youngs": E,
           "wall_s": wall, "ms_per_coarse_step": 1e3 * wall / steps,
           "mean_subiterations_per_fine_step": float(np.mean(its)), **obj,
           "dJ_mean_amplitude_dKB": obj["dJ_mean_amplitude"] * E / KB,
           "dJ_mean_deflection_dKB": obj["dJ_mean_deflection"] * E / KB}
    rep["_series"] = (tips, dtips)
    if check_steps and form == "fp64t":
        n = min(check_steps, steps)
        h = h_rel * E
        sched = its[:2 * n]
        up, _, its_u, _ = c5_tip_series("fp64", n, E + h, pin=sched)
        dn, _, its_d, _ = c5_tip_series("fp64", n, E - h, pin=sched)
        assert np.array_equal(its_u, sched) and np.array_equal(its_d, sched)
        # free (unpinned) stencil: the first fine step whose count differs
        fu = c5_tip_series("fp64", n, E + h)[2]
        fdn = c5_tip_series("fp64", n, E - h)[2]
        diff = [i for i in range(2 * n) if not (fu[i] == sched[i] == fdn[i])]
        rep["check_free_stencil_first_differing_fine_step"] = diff[0] if diff else None
        k = n
        fd = (up[:k] - dn[:k]) / (2 * h)
        scale = float(np.abs(dtips[:k]).max()) if k else 0.0
        tail = slice(k // 2, k)
        rep.update({"check_h_rel": h_rel, "check_steps": n,
                    "check_max_rel_diff_dtip": float(np.abs(dtips[:k] - fd).max() / scale) if scale else None,
                    "check_max_rel_diff_dtip_second_half": float(np.abs(dtips[tail] - fd[tail]).max() /
                                                                 np.abs(dtips[tail]).max()),
                    "check_dtip_scale": scale})
        if k:
            fo = c5_objectives(up[:k], None, y0, dhf)["J_mean_deflection_cells"]
            fm = c5_objectives(dn[:k], None, y0, dhf)["J_mean_deflection_cells"]
            to = c5_objectives(tips[:k], dtips[:k], y0, dhf)
            rep["check_dJ_mean_deflection_tangent"] = to["dJ_mean_deflection"]
            rep["check_dJ_mean_deflection_fd"] = (fo - fm) / (2 * h)
    return rep


def rel_l2(a, b, mask):
    d = (a - b)[mask]
    return float(np.sqrt((d * d).sum() / (b[mask] * b[mask]).sum()))


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("config", choices=["c1", "c3", "c4", "c5"])
    ap.add_argument("--precision", default="fp64,fp32s")
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--sample", type=int, default=20)
    ap.add_argument("--out", default=None)
    ap.add_argument("--forward", action="store_true",
                    help="c4: forward-step throughput with and without the sphere instead of the adjoint check")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="c4: time forward plus adjoint without the central differences")
    ap.add_argument("--every", type=int, default=200, help="c4 --adjoint-only: checkpoint interval")
    ap.add_argument("--grad", action="store_true",
                    help="c5: gradient of the mean tip amplitude with respect to bending stiffness")
    ap.add_argument("--check-steps", type=int, default=2000,
                    help="c5 --grad: coarse steps of the central-difference check (0: none)")
    ap.add_argument("--check-h", type=float, default=1e-3,
                    help="c5 --grad: relative step of the central differences in Young's modulus")
    a = ap.parse_args(argv)
    precs = a.precision.split(",")
    report = {"runs": []}
    if a.config == "c5" and a.grad:
        forms = [p for p in precs if p in ("fp64t", "fp32t")] or ["fp64t"]
        runs = {f: run_c5_grad(a.steps or 100000, a.check_steps, a.check_h, form=f) for f in forms}
        series = {f: r.pop("_series") for f, r in runs.items()}
        report = {"runs": list(runs.values())}
        if "fp64t" in series and "fp32t" in series:
            (t64, d64), (t32, d32) = series["fp64t"], series["fp32t"]
            dhf = c5_filament()[1].info["dh"]
            report["fp32t_vs_fp64t"] = {
                "tip_max_diff_cells": float(np.abs(t32 - t64).max() / dhf),
                "dtip_max_rel_diff": float(np.abs(d32 - d64).max() / np.abs(d64).max()),
                "dJ_mean_amplitude_dKB_rel_diff": abs(runs["fp32t"]["dJ_mean_amplitude_dKB"] -
                                                      runs["fp64t"]["dJ_mean_amplitude_dKB"]) /
                abs(runs["fp64t"]["dJ_mean_amplitude_dKB"])}
        print(json.dumps(report), flush=True)
        if a.out:
            with open(a.out, "w") as fh:
                json.dump(report, fh, indent=1)
        return 0
    if a.config == "c5":
        steps = a.steps or 100000
        series = {}
        for p in precs:
            r, tips = run_c5(p, steps)
            series[p] = tips
            report["runs"].append(r)
            print(json.dumps(r), flush=True)
        if "fp64" in series:
            dhf = c5_filament()[1].info["dh"]
            for p in precs:
                if p != "fp64":
                    report[f"tip_max_diff_cells_{p}_vs_fp64"] = float(np.abs(series[p] - series["fp64"]).max() / dhf)
        print(json.dumps({k: v for k, v in report.items() if k != "runs"}), flush=True)
        if a.out:
            with open(a.out, "w") as fh:
                json.dump(report, fh, indent=1)
        return 0
    if a.config == "c4" and a.forward:
        for p in precs:
            for body in (False, True):
                r = run_c4_forward(p, a.steps or 400, body)
                report["runs"].append(r)
                print(json.dumps(r), flush=True)
        if a.out:
            with open(a.out, "w") as fh:
                json.dump(report, fh, indent=1)
        return 0
    if a.config == "c4" and a.adjoint_only:
        for p in precs:
            r = run_c4_adjoint_timing(p, a.steps or 4000, a.every)
            report["runs"].append(r)
            print(json.dumps(r), flush=True)
        if a.out:
            with open(a.out, "w") as fh:
                json.dump(report, fh, indent=1)
        return 0
    if a.config == "c4":
        report = run_c4(tuple(precs), steps=a.steps or 4000)
        print(json.dumps(report), flush=True)
        if a.out:
            with open(a.out, "w") as fh:
                json.dump(report, fh, indent=1)
        return 0
    if a.config == "c1":
        steps = a.steps or 20000
        fields = {}
        for p in precs:
            r, (rho, u, case) = run_c1(p, steps)
            fields[p] = (rho, u)
            report["runs"].append(r)
            print(json.dumps(r), flush=True)
        if "fp64" in fields:
            m = case.lattyp != setups.SOLID
            for p in precs:
                if p != "fp64":
                    report[f"u_rel_l2_{p}_vs_fp64"] = rel_l2(fields[p][1], fields["fp64"][1], m)
    else:
        steps = a.steps or 100000
        marks = tuple(s for s in (1000, 10000, steps) if s <= steps)
        geo = c3_cylinder()
        res = {}
        for p in precs:
            r, extra = run_c3(p, steps, a.sample, fields_at=marks, geometry=geo)
            res[p] = extra
            report["runs"].append(r)
            print(json.dumps(r), flush=True)
        if "fp64" in res:
            parent, child = geo[0], geo[1]
            mp = ~np.isin(parent.lattyp, [setups.SOLID, setups.REFINED])
            mc = child.lattyp != setups.SOLID
            for p in precs:
                if p == "fp64":
                    continue
                errs = {}
                for s in marks:
                    (pr, pu), (cr, cu) = res[p]["snaps"][s]
                    (qr, qu), (dr, du) = res["fp64"]["snaps"][s]
                    errs[str(s)] = {"coarse": rel_l2(pu, qu, mp), "fine": rel_l2(cu, du, mc)}
                report[f"u_rel_l2_{p}_vs_fp64"] = errs
                a64 = next(r for r in report["runs"] if r["precision"] == "fp64")
                ap_ = next(r for r in report["runs"] if r["precision"] == p)
                report[f"{p}_vs_fp64_rel"] = {k: (abs(ap_[k] - a64[k]) / abs(a64[k]) if a64[k] else None)
                                              for k in ("Cd_mean", "Cl_amp", "St")}
                n = min(len(res[p]["cd"]), len(res["fp64"]["cd"]))
                report[f"cd_max_abs_diff_{p}_vs_fp64_first_1000_steps"] = float(
                    np.abs(res[p]["cd"][:1000 // a.sample] - res["fp64"]["cd"][:1000 // a.sample]).max())
                report[f"cd_series_len"] = n
    print(json.dumps({k: v for k, v in report.items() if k != "runs"}), flush=True)
    if a.out:
        with open(a.out, "w") as fh:
            json.dump(report, fh, indent=1)


if __name__ == "__main__":
    sys.exit(main())
