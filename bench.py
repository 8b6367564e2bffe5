#!/usr/bin/env python3
"""MLUPS and HBM-roofline fraction of the D3Q19 BGK 256^3 time step: FP32-shifted vs FP64.

Workload: D3Q19 BGK, 256^3 triply periodic Taylor-Green vortex (u0 = 0.05, k = 2 pi/256,
rho from the TGV pressure field, tau = 0.55), populations resident in HBM before timing.
A step is one fused AA-pattern stream-collide kernel over all 16,777,216 cells; the FP32-shifted
form is also timed with two steps per launch (temporal blocking: the local AA step and the
following neighbour step fused through LDS, reading and writing each population once per two
steps), and the faster of the two FP32-shifted runs is the reported value.

Timed: W warm-up steps, then exactly K steps between two stream synchronisations
(wall clock), with a HIP event pair around every launch on the grid's own stream (mean launch
duration).  Roofline of the reported kernel: the least HBM traffic one launch needs -- every
population read once and written once, 2*Q*s bytes per cell (304 B FP64, 152 B FP32-shifted,
76 B FP16-shifted), whether the launch advances one step or two -- over the mean launch
duration, against 8 TB/s.  For the two-step kernel that is the HBM bandwidth it uses; the
per-step figure (2*Q*s per cell per step, both steps counted) is reported beside it as
`cell_update_gbs`.  `traffic` is the HBM bytes per launch from rocprofv3's FETCH_SIZE and
WRITE_SIZE counters (profiles/tgv256_step_pmc.json, written by `bench.py --pmc-summary` from
two counter runs of this script), given only when that file was made from the kernel sources of
this tree.

Also reported: the same step in all three forms, the velocity L2 error and kinetic-energy
ratio of the shifted forms against FP64 after the configured 5,000 steps, and a CPU
baseline (the FP64 oracle restating LUMA's two-sweep AoS algorithm, OpenMP over x, on the
host's cores, on a bounded number of steps of the same 256^3 workload).

Prints one JSON line.  Only one GPU is used; --gpus above 1 exits with a message.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "MLUPS & %HBM roofline, D3Q19 256³: FP32-shifted vs FP64 at stated error"
PEAK_GBS = 8000.0            # MI355X HBM3E
ELEM = {"fp64": 8, "fp32s": 4, "fp16s": 2}
PMC_FILE = ROOT / "profiles" / "tgv256_step_pmc.json"
# the sources and build flags that decide the step kernels' code
KERNEL_SOURCES = ["luma_amd/csrc/aa_kernels.hpp", "luma_amd/csrc/fused_kernels.hpp", "luma_amd/csrc/lattice.hpp",
                  "luma_amd/csrc/launch.hpp", "luma_amd/csrc/dual.hpp", "luma_amd/csrc/lbm_f32.hip",
                  "luma_amd/csrc/lbm_f64.hip", "luma_amd/build.py"]


def kernel_source_hash() -> str:
    import hashlib
    h = hashlib.sha256()
    for rel in KERNEL_SOURCES:
        h.update(rel.encode())
        h.update((ROOT / rel).read_bytes())
    return h.hexdigest()[:16]


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def time_form(case, form, steps, warmup, two_step=None):
    """Time `steps` steps of one form; two_step = (shape, z chunk) runs them two per launch
    (temporal blocking), starting at odd parity so that every timed launch is a two-step one."""
    from luma_amd.grid import GpuGrid

    ncell = case.ncell
    with GpuGrid.from_case(case, form) as g:
        if two_step is not None:
            g.set_temporal_blocking(*two_step)
        g.run(warmup)
        if two_step is not None and g.t % 2 == 0:
            g.run(1)
        g.sync()
        t0 = time.perf_counter()
        _, ker_ms = g.run_timed(steps)
        g.sync()
        t1 = time.perf_counter()
    wall = t1 - t0
    bpc = 2 * case.Q * ELEM[form]
    spl = 2 if two_step is not None else 1
    launch_ms = spl * ker_ms               # run_timed reports per step; a launch is spl steps
    achieved = bpc * ncell / (launch_ms * 1e-3) / 1e9
    out = {
        "mlups": ncell * steps / wall / 1e6,
        "ms_per_step": wall * 1e3 / steps,
        "kernel_ms_mean": ker_ms,
        "launch_ms_mean": launch_ms,
        "steps_per_launch": spl,
        "bytes_per_cell_step": bpc,
        "hbm_min_bytes_per_launch": bpc * ncell,
        "achieved_gbs": achieved,
        "roofline_frac": achieved / PEAK_GBS,
        "cell_update_gbs": spl * achieved,
    }
    if two_step is not None:
        out["two_step"] = {"shape": two_step[0], "z_chunk": two_step[1]}
    return out


def slab_share(n, nz, steps, warmup, form="fp32s"):
    """One GPU's share of an 8-way z-sharded n^3 box: an n x n x nz slab plus two halo planes,
    the five face-crossing populations per face exchanged around every odd step (locally, the
    slab being its own neighbour), against the same slab stepped as a periodic box."""
    import dataclasses

    import numpy as np

    from luma_amd import setups
    from luma_amd.grid import GpuGrid

    box = setups.taylor_green((n, n, nz), u0=0.05, tau=0.55, dims=3)
    pad = lambda a: np.concatenate([a[:, :, -1:], a, a[:, :, :1]], axis=2)
    slab = dataclasses.replace(box, K=nz + 2, rho=pad(box.rho), u=pad(box.u))
    out = {"shape": [n, n, nz], "form": form}
    for key, case, zs in (("periodic", box, False), ("slab", slab, True)):
        with GpuGrid.from_case(case, form) as g:
            if zs:
                g.set_zslab(True)
            g.run(warmup)
            g.sync()
            tot_ms, ker_ms = g.run_timed(steps)
        out[key] = {"step_us": 1e3 * tot_ms / steps, "kernel_us": 1e3 * ker_ms}
    out["halo_us_per_step"] = out["slab"]["step_us"] - out["slab"]["kernel_us"]
    out["face_bytes_per_step"] = 5 * n * n * ELEM[form]
    out["mlups"] = n * n * nz / out["slab"]["step_us"]
    return out


def error_run(case, nsteps, two_step=None):
    """Run every form nsteps from the same state; u error and KE of shifted forms vs FP64.
    two_step = (shape, z chunk): the FP32-shifted run uses the two-steps-per-launch kernel, the
    path whose throughput is reported."""
    import numpy as np

    from luma_amd.grid import GpuGrid

    out = {}
    with GpuGrid.from_case(case, "fp64") as g64:
        ke0 = g64.moments()["ke"]
        g64.run(nsteps - 1)
        g64.set_store_macro(True)
        g64.run(1)
        m64 = g64.moments()
        _, u64 = g64.download_macro()
    out["fp64"] = {"ke_ratio": m64["ke"] / ke0}
    norm = np.linalg.norm(u64)
    for form in ("fp32s", "fp16s"):
        with GpuGrid.from_case(case, form) as g:
            if form == "fp32s" and two_step is not None:
                g.set_temporal_blocking(*two_step)
            g.run(nsteps - 1)
            g.set_store_macro(True)
            g.run(1)
            m = g.moments()
            _, u = g.download_macro()
        out[form] = {"u_l2_rel_vs_fp64": float(np.linalg.norm(u - u64) / norm),
                     "ke_ratio": m["ke"] / ke0,
                     "ke_rel_vs_fp64": m["ke"] / m64["ke"] - 1.0}
        del u
    # linear (Stokes) decay of the TGV mode, |k|^2 = 3 k^2; the 3D vortex at this Reynolds
    # number also cascades energy to small scales, so the measured ratio lies below it
    nu = (0.55 - 0.5) / 3.0
    k2 = 3 * (2 * math.pi / case.N) ** 2
    out["ke_ratio_stokes_decay"] = math.exp(-2 * nu * k2 * nsteps)
    return out


def cpu_baseline(case, target_s=15.0, warmup=5):
    from oracle.oracle import OracleGrid

    threads = min(16, len(os.sched_getaffinity(0)))
    g = OracleGrid(case.dims, case.Q, case.N, case.M, case.K, omega=case.omega)
    g.init_feq(case.rho, case.u)
    t0 = time.perf_counter()
    g.run(warmup, threads)
    one = (time.perf_counter() - t0) / warmup
    n = int(max(2, min(50, target_s / max(one, 1e-3))))
    t0 = time.perf_counter()
    c0 = time.process_time()
    g.run(n, threads)
    dt = time.perf_counter() - t0
    cpu_s = time.process_time() - c0
    g.close()
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            model = next((ln.split(":", 1)[1].strip() for ln in fh if ln.startswith("model name")), None)
    except OSError:
        pass
    # LUMA's own figure (main_lbm.cpp:455-457): GridManager::activeCellOps (GridManager.cpp:642-679,
    # the level-0 cell count on one level) over the outer-loop time in ms, times 1000, where the
    # outer-loop time comes from clock() (process CPU time, summed over threads) per step
    ops = case.ncell
    luma_fig = ops / ((cpu_s / n) * 1e3 * 1000)
    return {"value": case.ncell * n / dt / 1e6, "unit": "MLUPS", "cores": threads, "kind": "port",
            "luma_active_cell_ops_per_time": luma_fig,
            "host": {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)), "cpu_model": model,
                     "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS")},
            "sample": f"D3Q19 BGK TGV {case.N}^3 FP64, {n} steps after {warmup} warm-up steps: oracle restating "
                      f"LUMA's LBM_multi_opt (AoS f/fNew pull, two sweeps), OpenMP over x; value from the "
                      f"wall clock, luma_active_cell_ops_per_time as LUMA's log computes it (CPU time)"}


def pmc_summary(fetch_dir: str, write_dir: str, out_path: str, n: int = 256) -> dict:
    """HBM bytes per launch of the step kernels from two rocprofv3 counter runs of this script
    (one with --pmc FETCH_SIZE, one with --pmc WRITE_SIZE; counters in KiB per dispatch).  Each
    counter is scaled by what it reads for a kernel of known traffic in the same run: k_moments
    reads every population slot once (19*s B/cell), k_init_feq writes every slot once."""
    import csv
    import re

    def load(d, counter):
        vals = {}
        for f in sorted(Path(d).rglob("*counter_collection.csv")):
            with open(f, newline="") as fh:
                for row in csv.DictReader(fh):
                    if row.get("Counter_Name") == counter:
                        vals.setdefault(row["Kernel_Name"], []).append(float(row["Counter_Value"]) * 1024.0)
        if not vals:
            raise SystemExit(f"no {counter} rows under {d}")
        return vals

    fetch, write = load(fetch_dir, "FETCH_SIZE"), load(write_dir, "WRITE_SIZE")
    cells = n ** 3
    forms = {"fp64": "P64", "fp32s": "P32S", "fp16s": "P16S"}

    def mean_of(vals, pattern):
        xs = [v for k, vs in vals.items() if re.search(pattern, k) for v in vs]
        return sum(xs) / len(xs) if xs else None

    def factor(vals, kernel):
        fs = []
        for form, P in forms.items():
            m = mean_of(vals, rf"{kernel}<19, (lumagpu::)?{P}\b")
            if m is not None:
                fs.append(m / (19 * ELEM[form] * cells))
        if not fs:
            raise SystemExit(f"no {kernel} dispatches to calibrate against")
        return sum(fs) / len(fs)

    ff, wf = factor(fetch, "k_moments"), factor(write, "k_init_feq")
    out = {"kernel_source_hash": kernel_source_hash(), "workload": f"D3Q19 TGV {n}^3", "fetch_factor": round(ff, 4),
           "write_factor": round(wf, 4),
           "method": "FETCH_SIZE / fetch_factor + WRITE_SIZE / write_factor, means over dispatches"}
    for form, P in forms.items():
        pat = rf"k_aa_periodic<19, (lumagpu::)?{P}, (true|false), false"
        fm, wm = mean_of(fetch, pat), mean_of(write, pat)
        if fm is not None and wm is not None:
            out[form] = {"hbm_bytes_per_launch": round(fm / ff + wm / wf), "min_bytes_per_launch": 2 * 19 * ELEM[form] * cells,
                         "steps_per_launch": 1}
    pat = r"k_aa_fused2<(lumagpu::)?P32S"
    fm, wm = mean_of(fetch, pat), mean_of(write, pat)
    if fm is not None and wm is not None:
        kn = sorted(k for k in fetch if re.search(pat, k))
        out["fp32s_2step"] = {"hbm_bytes_per_launch": round(fm / ff + wm / wf), "min_bytes_per_launch": 2 * 19 * 4 * cells,
                              "steps_per_launch": 2, "kernels": kn}
    Path(out_path).write_text(json.dumps(out, indent=1) + "\n")
    return out


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "--pmc-summary":
        if len(sys.argv) != 5:
            sys.exit("usage: bench.py --pmc-summary FETCH_DIR WRITE_DIR OUT_JSON")
        print(json.dumps(pmc_summary(*sys.argv[2:5])))
        return
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=256, help="cube edge (cells)")
    ap.add_argument("--error-steps", type=int, default=5000, help="0 skips the error run")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--slab-steps", type=int, default=200, help="0 skips the per-GPU slab measurement")
    ap.add_argument("--forms", default="fp64,fp32s,fp16s")
    ap.add_argument("--two-step", default="24,128",
                    help="FP32-shifted two-steps-per-launch kernel: tile shape code, z planes per block "
                         "('' skips it)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported in this build "
              f"(the multi-GPU slab decomposition is a design, not a measured path)", file=sys.stderr)
        sys.exit(2)

    from luma_amd import setups

    case = setups.taylor_green(args.n, u0=0.05, tau=0.55, dims=3)
    forms = [f for f in args.forms.split(",") if f]
    res = {}
    for form in forms:
        log(f"[bench] timing {form} {args.n}^3: warmup {args.warmup}, steps {args.steps}")
        res[form] = time_form(case, form, args.steps, args.warmup)
        log(f"[bench] {form}: {res[form]['mlups']:.0f} MLUPS, kernel {res[form]['kernel_ms_mean']:.3f} ms, "
            f"{res[form]['achieved_gbs']:.0f} GB/s ({100 * res[form]['roofline_frac']:.1f}% of HBM)")
    if args.two_step and "fp32s" in forms:
        shape, zc = (int(x) for x in args.two_step.split(","))
        log(f"[bench] timing fp32s {args.n}^3 two steps per launch (shape {shape}, z chunk {zc})")
        from luma_amd._abi import LumaGpuError
        try:
            res["fp32s_2step"] = time_form(case, "fp32s", args.steps, args.warmup, two_step=(shape, zc))
        except LumaGpuError as e:
            log(f"[bench] two-step kernel not available for this grid: {e}")
        if "fp32s_2step" in res:
            r2 = res["fp32s_2step"]
            log(f"[bench] fp32s_2step: {r2['mlups']:.0f} MLUPS, launch {r2['launch_ms_mean']:.3f} ms "
                f"per two steps, {r2['achieved_gbs']:.0f} GB/s of HBM ({100 * r2['roofline_frac']:.1f}% of 8 TB/s), "
                f"{r2['cell_update_gbs']:.0f} GB/s in per-step bytes")
    err = None
    if args.error_steps > 0 and "fp64" in forms:
        log(f"[bench] error run: {args.error_steps} steps per form")
        two = None
        if "fp32s_2step" in res:
            ts = res["fp32s_2step"]["two_step"]
            two = (ts["shape"], ts["z_chunk"])
        err = error_run(case, args.error_steps, two_step=two)
        if two is not None:
            err["fp32s"]["kernel"] = "two steps per launch"
        log(f"[bench] errors: {json.dumps(err)}")
    slab = None
    if args.slab_steps > 0:
        log(f"[bench] one GPU's share of an 8-way z-sharded {args.n}^3 box")
        slab = slab_share(args.n, args.n // 8, args.slab_steps, args.warmup)
        log(f"[bench] slab: {json.dumps(slab)}")
    cpu = None
    if not args.no_cpu:
        log("[bench] CPU baseline")
        cpu = cpu_baseline(case)
        log(f"[bench] cpu: {cpu['value']:.1f} MLUPS on {cpu['cores']} threads")

    head = "fp32s" if "fp32s" in res else forms[0]
    if "fp32s_2step" in res and res["fp32s_2step"]["mlups"] > res["fp32s"]["mlups"]:
        head = "fp32s_2step"
    form = head.split("_")[0]
    r = res[head]
    # HBM bytes per launch of this kernel from the rocprofv3 counter runs, if they were made from
    # this tree's kernel sources
    traffic, stale = None, None
    if PMC_FILE.exists():
        pmc = json.loads(PMC_FILE.read_text())
        stale = pmc.get("kernel_source_hash") != kernel_source_hash()
        if not stale:
            traffic = pmc.get(head, {}).get("hbm_bytes_per_launch")
    steps_per_launch = r["steps_per_launch"]
    launch_ms = r["launch_ms_mean"]
    line = {
        "metric": METRIC,
        "value": round(r["mlups"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if form != "fp64" else "f64",
        "data": "synthetic",
        "config": {"workload": f"D3Q19 BGK Taylor-Green {args.n}^3 triply periodic, u0=0.05, tau=0.55",
                   "form": form, "cells": case.ncell, "storage": {"fp64": "f64", "fp32s": "f32 shifted (f - w rho0)",
                                                                  "fp16s": "f16 shifted (f - w rho0)"}[form],
                   "kernel": "two AA steps per launch (temporal blocking)" if steps_per_launch == 2
                             else "one AA step per launch",
                   "bytes_per_cell_step": r["bytes_per_cell_step"], "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(r["achieved_gbs"], 1), "peak": PEAK_GBS, "unit": "GB/s",
                     "frac": round(r["roofline_frac"], 4), "traffic": traffic,
                     "traffic_over_min": None if traffic is None else round(traffic / r["hbm_min_bytes_per_launch"], 4),
                     "traffic_stale": stale, "min_bytes_per_launch": r["hbm_min_bytes_per_launch"],
                     "steps_per_launch": steps_per_launch, "launch_ms": round(launch_ms, 5),
                     "cell_update_gbs": round(r["cell_update_gbs"], 1)},
        "cpu_baseline": cpu,
        "forms": {k: {kk: (round(vv, 5) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                  for k, v in res.items()},
    }
    if "fp64" in res and "fp32s" in res:
        line["fp32s_over_fp64"] = round(r["mlups"] / res["fp64"]["mlups"], 3) if form == "fp32s" else None
    if err is not None:
        line["error"] = {"steps": args.error_steps, **err}
    if slab is not None:
        line["slab_share"] = {k: (round(v, 3) if isinstance(v, float) else v) for k, v in slab.items()}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
