#!/usr/bin/env python3
"""Converged modes/s of batched solve_k on one MI355X.

Workload (one step): solve_k for the 8 lowest modes of an x-cut thin-film LiNbO3 rib
(diag eps = (4.89, 4.57, 4.89), top width 1.2 um, 0.6 um film, 0.3 um etch, 60 deg
sidewalls, SiO2 below, air above) in an 8 x 4 um cell on a 256 x 256 grid, omega =
1/1.55 um^-1, Kottke-smoothed eps^-1 (synthetic, generated in-process; resident on the
GPU before timing).  FP32 operator phase + FP64 refinement phase, eigenvectors left on
the device.  A mode counts as converged when its omega^2 relative error, evaluated
against the FP64 oracle's k through the oracle's slope d(omega^2)/dk, is <= 1e-8.

Prints one JSON line (see the build contract): value = converged modes / s.
"""
import argparse
import hashlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "optimode.jl_amd"))

METRIC = "converged modes/s at 256² (ω² rel err ≤1e-8); LOBPCG iters vs plain"
HBM_PEAK_GBS = 8000.0
OMEGA = 1 / 1.55
NEV = 8
REL_TOL = 1e-8
# HBM bytes per algorithmic byte of the FP32 operator application on this workload, from
# rocprofv3 FETCH_SIZE + WRITE_SIZE of k_rpass1/2/3<float, OP_M> (FETCH scaled by 2.0, WRITE
# by 1.0: the factors a 256 MiB copy kernel reads back on the same box); the measurement is
# kept in profiles/solve_k_tfln256_pmc_traffic.md.  Update when the kernels change.
OP_TRAFFIC_RATIO = 3.82


def cpu_baseline(seconds_cap=60.0):
    """The FP64 oracle (NumPy + multithreaded pocketfft) timed on the host: solve_k for the
    fundamental mode (nev = 1) of the same structure."""
    import numpy as np
    try:
        ncores = len(os.sched_getaffinity(0))
    except AttributeError:
        ncores = os.cpu_count() or 1
    ncores = max(1, min(16, ncores))
    os.environ["ORACLE_FFT_WORKERS"] = str(ncores)
    import oracle as O  # the checker: timed here as the CPU baseline only
    from omx import geometry as G
    g = O.Grid(8.0, 4.0, 256, 256)
    ei = G.tfln_rib(g)
    t = time.perf_counter()
    km, _ = O.solve_k(OMEGA, ei, g, nev=1, k_tol=1e-10, eig_tol=1e-8, maxiter=300, lock_band_only=True)
    dt = time.perf_counter() - t
    return {"value": 1.0 / dt, "unit": "modes/s", "cores": ncores, "kind": "port",
            "sample": f"FP64 NumPy oracle solve_k, TFLN rib 256x256, nev=1 (fundamental), eig_tol 1e-8, "
                      f"k_tol 1e-10: {dt:.1f} s for 1 mode (k={km[0]:.8f})"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--no-plain", action="store_true", help="skip the plain-algorithm comparison solve")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}): every solve here fits one MI355X and "
              "multi-GPU sharding is by independent instances, which this build does not launch.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import omx
    from omx import geometry as G

    grid = omx.Grid(8.0, 4.0, 256, 256)
    ei = G.tfln_rib(grid)
    gold = np.load(os.path.join(ROOT, "tests", "golden", "c2_tfln_256.npz"))
    sha = hashlib.sha256(np.ascontiguousarray(ei, dtype=np.float64).tobytes()).hexdigest()
    if sha != str(gold["epsinv_sha256"]):
        raise RuntimeError("the generated eps^-1 does not match the one the golden k values were computed for")
    k_ref, slope_ref = gold["kmags"], gold["dw2dk"]

    solver = omx.HIPSolver(precision=a.precision, profile=True)
    kw = dict(nev=NEV, k_tol=1e-10, eig_tol=1e-8, max_eigsolves=60, maxiter=200, fetch_evecs=False,
              return_info=True)
    for _ in range(a.warmup):
        omx.solve_k(OMEGA, ei, grid, solver, **kw)
    # reset the live kernel statistics so that they cover the timed region only
    h = solver.handle(grid, ei)
    omx._lib.check(omx._lib.lib().omx_set_option(h.h, b"profile", 1.0), h.h)

    converged = 0
    infos = []
    errs = []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        km, _, info = omx.solve_k(OMEGA, ei, grid, solver, **kw)
        rel = np.abs(slope_ref * (km - k_ref)) / OMEGA ** 2
        converged += int(np.sum(rel <= REL_TOL))
        errs.append(float(rel.max()))
        infos.append(info)
    t1 = time.perf_counter()
    elapsed = t1 - t0
    ms_per_step = 1e3 * elapsed / a.steps
    value = converged / elapsed

    op_ms, op_n, op_bytes = omx.kernel_stats(solver, "op_fp32")

    # the reference's algorithm on the same kernels (FP64, nev-column block per band, random
    # start per band, exact eigensolves): one warm-up and one timed solve, outside the timed
    # region above
    plain = None
    if not a.no_plain:
        sp = omx.HIPSolver(precision="fp64", plain=True)
        omx.solve_k(OMEGA, ei, grid, sp, **kw)
        tp = time.perf_counter()
        kmp, _, infop = omx.solve_k(OMEGA, ei, grid, sp, **kw)
        dtp = time.perf_counter() - tp
        relp = np.abs(slope_ref * (kmp - k_ref)) / OMEGA ** 2
        nconv = int(np.sum(relp <= REL_TOL))
        plain = {"modes_per_s": round(nconv / dtp, 3), "ms_per_step": round(1e3 * dtp, 2),
                 "modes_converged": nconv, "lobpcg_iters_fp64": infop["lobpcg_iters_fp64"],
                 "newton_steps": infop["n_newton"], "matvecs": infop["n_matvec"],
                 "speedup": round(value / (nconv / dtp), 2) if nconv else None}
    roofline = None
    if op_n > 0 and op_ms > 0:
        achieved = op_bytes / (op_ms * 1e-3) / 1e9
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": round(OP_TRAFFIC_RATIO * op_bytes / op_n),
                    "kernel": "FP32 Helmholtz operator application (3 fused passes)",
                    "launches": op_n, "avg_us": round(1e3 * op_ms / op_n, 2),
                    "avg_alg_bytes": round(op_bytes / op_n)}

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "modes/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64",
        "data": "synthetic",
        "config": {
            "workload": "solve_k, x-cut TFLN rib (diag eps 4.89/4.57/4.89, 1.2 um top, 0.6 um film, 0.3 um etch, "
                        "60 deg), SiO2/air, 8x4 um, 256x256, omega=1/1.55, 8 modes, k_tol 1e-10, eig_tol 1e-8",
            "grid": [256, 256],
            "nev": NEV,
            "precision": a.precision,
            "modes_converged": converged,
            "modes_total": NEV * a.steps,
            "max_rel_err_omega2": max(errs),
            "lobpcg_iters_fp32_per_step": float(np.mean([i["lobpcg_iters_fp32"] for i in infos])),
            "lobpcg_iters_fp64_per_step": float(np.mean([i["lobpcg_iters_fp64"] for i in infos])),
            "newton_steps_per_step": float(np.mean([i["n_newton"] for i in infos])),
            "matvecs_per_step": float(np.mean([i["n_matvec"] for i in infos])),
            "parallelism": "single GPU; bands batched as independent Newton instances",
        },
        "plain": plain,
        "roofline": roofline,
        "cpu_baseline": None if a.no_cpu_baseline else cpu_baseline(),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
