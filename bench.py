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
# rocprofv3 FETCH_SIZE + WRITE_SIZE of the k_rpass1/2/3<float, OP_M> launches of the
# refinement's correction solves (8 vectors, all active, so their algorithmic bytes are
# exact), FETCH scaled by 2.0 and WRITE by 1.0 (the factors `bench.py --workload copy`, a
# 256 MiB device copy, reads back on the same box); the measurement is kept in
# profiles/solve_k_tfln256_pmc_traffic.md.  Update when the kernels change.
OP_TRAFFIC_RATIO = 3.78


def cpu_baseline(seconds_cap=60.0):
    """The FP64 oracle (NumPy + pocketfft restatement of the reference's CPU path) timed on the
    host: solve_k for the fundamental mode (nev = 1) of the same structure on all cores (up to
    16) and on one core (the reference runs FFTW single-threaded, src/MaxwellEigenmodes.jl:47),
    and the 128^2 rectangle configuration (4 modes) on all cores.  The reference's own
    published 256^2 figure is quoted beside them."""
    import numpy as np
    from threadpoolctl import threadpool_limits
    try:
        ncores = len(os.sched_getaffinity(0))
    except AttributeError:
        ncores = os.cpu_count() or 1
    ncores = max(1, min(16, ncores))
    os.environ["ORACLE_FFT_WORKERS"] = str(ncores)
    import oracle as O  # the oracle, timed here as the CPU baseline only
    import oracle.maxwell as OM
    from omx import geometry as G
    g = O.Grid(8.0, 4.0, 256, 256)
    ei = G.tfln_rib(g)

    def timed(grid, eps, nev, cores):
        OM._WORKERS = cores
        with threadpool_limits(limits=cores):
            t = time.perf_counter()
            km, _ = O.solve_k(OMEGA, eps, grid, nev=nev, k_tol=1e-10, eig_tol=1e-8, maxiter=300, lock_band_only=True)
            return time.perf_counter() - t, km

    dt, km = timed(g, ei, 1, ncores)
    dt1, _ = timed(g, ei, 1, 1)
    g1 = O.Grid(6.0, 4.0, 128, 128)
    dtc1, kc1 = timed(g1, G.rect_waveguide(g1), 4, ncores)
    OM._WORKERS = ncores
    return {"value": 1.0 / dt, "unit": "modes/s", "cores": ncores, "kind": "port",
            "sample": f"FP64 NumPy oracle solve_k, TFLN rib 256x256, nev=1 (fundamental), eig_tol 1e-8, "
                      f"k_tol 1e-10: {dt:.1f} s for 1 mode (k={km[0]:.8f}) on {ncores} cores",
            "one_core": {"value": 1.0 / dt1, "unit": "modes/s", "cores": 1,
                         "sample": f"the same solve on 1 core: {dt1:.1f} s"},
            "rect128_4modes": {"value": 4.0 / dtc1, "unit": "modes/s", "cores": ncores,
                               "sample": f"Si3N4 rectangle 128x128, nev=4, eig_tol 1e-8, k_tol 1e-10: {dtc1:.1f} s"},
            "reference_published": "README.md:363 (reference): solve_k 256x256, nev=1, KrylovKit Float64 on a "
                                   "4-core container: 125 s (0.008 modes/s); Gaussian-bump index profile"}


EME_SLICES = 200      # cross-sections of the whole coupler
EME_LENGTH = 200.0    # um


def eme_cell(grid, i, backend="gpu"):
    """Cross-section i of the adiabatic Si3N4 coupler: gap 0.3 -> 1.5 um and widths
    0.8 <-> 1.2 um linear along the device, 0.8 um thick, SiO2 cladding."""
    from omx import geometry as G
    s = i / (EME_SLICES - 1)
    return G.coupler_slice(grid, 0.3 + 1.2 * s, 0.8 + 0.4 * s, 1.2 - 0.4 * s, h=0.8, backend=backend)


def eme_workload(a):
    """EME slice chain at 512 x 256 with 16 modes per slice: per step (one slice) GPU
    smoothing, solve_k warm-started from the previous slice, the power-normalised mode set
    and the overlaps with the previous slice on the GPU, and the host interface/propagation
    star products.  Prints one JSON line (slices/s and modes/s)."""
    import numpy as np
    import omx
    from omx import eme as X

    grid = omx.Grid(8.0, 4.0, 512, 256)
    nev = 16
    L = EME_LENGTH / EME_SLICES
    kw = dict(k_tol=1e-10, eig_tol=1e-8)
    nw, ns = a.warmup, a.steps
    cells = [(lambda i=i: eme_cell(grid, i)) for i in range(nw + ns)]
    solver = omx.HIPSolver(precision=a.precision)
    # the coupler network starts the chain's cold first slice (and any cold restart)
    from omx import warmstart as WS
    learned = WS.load_coupler() if a.start != "plain" else None
    # warm-up slices (also the chain's start), then the timed slices continue the same chain
    X.eme(cells[:nw], [L] * nw, OMEGA, grid, solver, nev=nev, warmstart=True, learned=learned, **kw)
    t0 = time.perf_counter()
    res = X.eme(cells[nw:], [L] * ns, OMEGA, grid, solver, nev=nev, warmstart=True, continue_chain=True, **kw)
    elapsed = time.perf_counter() - t0
    conv = sum(int(st["converged"]) for st in res.stats)
    # the last timed slice against a cold FP64 solve at a tighter tolerance
    ref = omx.HIPSolver(precision="fp64")
    ei_last = eme_cell(grid, nw + ns - 1)
    kr, _ = omx.solve_k(OMEGA, ei_last, grid, ref, nev=nev, k_tol=1e-12, eig_tol=1e-10, fetch_evecs=False)
    rel = float(np.max(np.abs(res.kmags[-1] ** 2 - kr ** 2) / kr ** 2))
    # a cold chain over the first few timed slices for the warm-start gain
    cold_n = min(ns, 4)
    sc = omx.HIPSolver(precision=a.precision)
    tc = time.perf_counter()
    rc = X.eme(cells[nw:nw + cold_n], [L] * cold_n, OMEGA, grid, sc, nev=nev, warmstart=False, **kw)
    cold_ms = 1e3 * (time.perf_counter() - tc) / cold_n
    # the same cold slices from the coupler network's guess
    cold_learned = None
    if learned is not None:
        tl = time.perf_counter()
        rl = X.eme(cells[nw:nw + cold_n], [L] * cold_n, OMEGA, grid, sc, nev=nev, warmstart=False, learned=learned,
                   **kw)
        cold_learned = dict(ms_per_slice=round(1e3 * (time.perf_counter() - tl) / cold_n, 2),
                            newton_per_slice=float(np.mean([st["n_newton"] for st in rl.stats])),
                            lobpcg_iters_fp32_per_slice=float(np.mean([st["lobpcg_iters_fp32"] for st in rl.stats])),
                            max_rel_diff_k_vs_plain_cold=float(max(np.max(np.abs(x / y - 1))
                                                                   for x, y in zip(rl.kmags, rc.kmags))))
    out = {
        "metric": "EME slices/s at 512x256, 16 modes per slice (previous-slice warm start)",
        "value": round(ns / elapsed, 3),
        "unit": "slices/s",
        "n_gpus": 1,
        "steps": ns,
        "warmup": nw,
        "ms_per_step": round(1e3 * elapsed / ns, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64" if a.precision == "mixed" else "fp64",
        "data": "synthetic",
        "config": {
            "workload": "EME chain: slices of an adiabatic Si3N4 coupler (gap 0.3->1.5 um, widths 0.8<->1.2 um, "
                        "0.8 um thick, SiO2), 8x4 um, 512x256, omega=1/1.55, 16 modes/slice, k_tol 1e-10, "
                        "eig_tol 1e-8; per slice: GPU smoothing, warm solve_k, mode set, overlaps, star products",
            "modes_per_s": round(nev * conv / elapsed, 2),
            "slices_converged": conv,
            "newton_per_slice": float(np.mean([st["n_newton"] for st in res.stats])),
            "lobpcg_iters_fp32_per_slice": float(np.mean([st["lobpcg_iters_fp32"] for st in res.stats])),
            "fp64_rr_steps_per_slice": float(np.mean([st["lobpcg_iters_fp64"] for st in res.stats])),
            "last_slice_max_rel_err_k2_vs_cold_fp64": rel,
            "cold_chain_ms_per_slice": round(cold_ms, 2),
            "cold_newton_per_slice": float(np.mean([st["n_newton"] for st in rc.stats])),
            "cold_lobpcg_iters_fp32_per_slice": float(np.mean([st["lobpcg_iters_fp32"] for st in rc.stats])),
            "cold_learned_start": cold_learned,
            "first_slice_start": "learned (omx/data/warmnet_coupler512x256.npz)" if learned is not None else "plain",
            "power_coupling_1_1": X.power_coupling(res),
        },
    }
    print(json.dumps(out), flush=True)


PER_GRID = (4.0, 3.0, 0.45, 128, 96, 24)   # Dx, Dy, period Dz (um), Nx, Ny, Nz
PER_OMEGA = 0.5
PER_NEV = 4


def periodic_cpu_baseline(ncores):
    """The FP64 oracle on the host: the fundamental Bloch mode of the same 3D cell."""
    os.environ["ORACLE_FFT_WORKERS"] = str(ncores)
    import oracle as O  # the oracle, timed here as the CPU baseline only
    from oracle import periodic as P
    from omx import geometry as G
    g = P.Grid3(*PER_GRID)
    ei = G.bragg_grating(g)
    t = time.perf_counter()
    km, _ = O.solve_k(PER_OMEGA, ei, g, nev=1, k_tol=1e-10, eig_tol=1e-8, maxiter=300)
    dt = time.perf_counter() - t
    return {"value": 1.0 / dt, "unit": "modes/s", "cores": ncores, "kind": "port",
            "sample": f"FP64 NumPy oracle solve_k on the same 128x96x24 cell, nev=1 (fundamental), eig_tol 1e-8, "
                      f"k_tol 1e-10: {dt:.1f} s for 1 mode (k={km[0]:.8f})"}


def periodic_workload(a):
    """solve_k on one period of a 3D Bragg-grating waveguide (128 x 96 x 24 voxels, period
    0.45 um, 4 Bloch modes): the z-periodic path (three-kernel yz-slab 3D operator).  Prints one JSON
    line with modes/s, the FP32 3D operator's roofline and the CPU baseline."""
    import numpy as np
    import omx
    from omx import geometry as G

    grid = omx.Grid3(*PER_GRID)
    ei = G.bragg_grating(grid)
    solver = omx.HIPSolver(precision=a.precision, profile=True)
    kw = dict(nev=PER_NEV, k_tol=1e-10, eig_tol=1e-8, max_eigsolves=60, maxiter=300, fetch_evecs=False,
              return_info=True)
    for _ in range(a.warmup):
        omx.solve_k(PER_OMEGA, ei, grid, solver, **kw)
    h = solver.handle(grid, ei)
    omx._lib.check(omx._lib.lib().omx_set_option(h.h, b"profile", 1.0), h.h)
    infos, kms = [], []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        km, _, info = omx.solve_k(PER_OMEGA, ei, grid, solver, **kw)
        kms.append(km)
        infos.append(info)
    elapsed = time.perf_counter() - t0
    op_ms, op_n, op_bytes = omx.kernel_stats(solver, "op_fp32" if a.precision == "mixed" else "op_fp64")
    # accuracy: against a cold FP64 solve at tighter tolerances, |Δω²| = slope |Δk|
    ref = omx.HIPSolver(precision="fp64")
    kr, evr = omx.solve_k(PER_OMEGA, ei, grid, ref, nev=PER_NEV, k_tol=1e-12, eig_tol=1e-10, maxiter=300)
    dk, dL = omx.period_derivatives(kr, evr, ei, grid, ref)
    errs = [float(np.max(np.abs(dk * (km - kr)) / PER_OMEGA ** 2)) for km in kms]
    conv = sum(int(np.sum(np.abs(dk * (km - kr)) / PER_OMEGA ** 2 <= REL_TOL)) for km in kms)
    roofline = None
    if op_n > 0 and op_ms > 0:
        achieved = op_bytes / (op_ms * 1e-3) / 1e9
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                    "kernel": "3D Helmholtz operator application (x rows, yz slabs in LDS, x rows)",
                    "launches": op_n, "avg_us": round(1e3 * op_ms / op_n, 2), "avg_alg_bytes": round(op_bytes / op_n)}
    ncores = 1
    try:
        ncores = max(1, min(16, len(os.sched_getaffinity(0))))
    except AttributeError:
        ncores = max(1, min(16, os.cpu_count() or 1))
    out = {
        "metric": "converged Bloch modes/s, 3D Bragg-grating waveguide 128x96x24 (omega^2 rel err <=1e-8)",
        "value": round(conv / elapsed, 3),
        "unit": "modes/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(1e3 * elapsed / a.steps, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64" if a.precision == "mixed" else "fp64",
        "data": "synthetic",
        "config": {
            "workload": "solve_k_periodic: one period (0.45 um) of a Gaussian-core Bragg grating (eps 4.0 core, "
                        "2.1 background, width 1.0 um modulated +-50%, height 0.4 um), 4x3 um cross-section, "
                        "128x96x24 voxels, omega=0.5, 4 modes, k_tol 1e-10, eig_tol 1e-8",
            "grid": list(PER_GRID[3:]),
            "nev": PER_NEV,
            "precision": a.precision,
            "kmags": [float(x) for x in kr],
            "dw2_dperiod": [float(x) for x in dL],
            "modes_converged": conv,
            "modes_total": PER_NEV * a.steps,
            "max_rel_err_omega2_vs_fp64": max(errs),
            "lobpcg_iters_fp32_per_step": float(np.mean([i["lobpcg_iters_fp32"] for i in infos])),
            "fp64_rr_steps_per_step": float(np.mean([i["lobpcg_iters_fp64"] for i in infos])),
            "newton_steps_per_step": float(np.mean([i["n_newton"] for i in infos])),
            "matvecs_per_step": float(np.mean([i["n_matvec"] for i in infos])),
        },
        "roofline": roofline,
        "cpu_baseline": None if a.no_cpu_baseline else periodic_cpu_baseline(ncores),
    }
    print(json.dumps(out), flush=True)


def batch_problems(seed=3, n=64):
    """The mixed-grid batch: seeded ridges (w ~ U(0.8, 2.5) um, h ~ U(0.3, 0.9) um,
    sidewall ~ U(60, 90) deg, core eps ~ U(3.5, 4.9)) on SiO2 under air, grids uniform in
    {128^2, 192^2, 256^2, 384^2} over an 8 x 4 um cell, omega ~ U(0.5, 1.0), nev uniform in
    4..16; full 3x3 smoothed tensors (xy terms at the slanted faces), smoothed on the GPU."""
    import numpy as np
    import omx
    from omx import geometry as G
    rng = np.random.default_rng(seed)
    probs = []
    for _ in range(n):
        N = int(rng.choice([128, 192, 256, 384]))
        g = omx.Grid(8.0, 4.0, N, N)
        w, h, sw, ec = rng.uniform(0.8, 2.5), rng.uniform(0.3, 0.9), rng.uniform(60, 90), rng.uniform(3.5, 4.9)
        om, nev = rng.uniform(0.5, 1.0), int(rng.integers(4, 17))
        probs.append(dict(grid=g, epsinv=G.ridge_waveguide(g, w, h, sw, ec, backend="gpu"), omega=om, nev=nev,
                          k_tol=1e-10, eig_tol=1e-8, max_eigsolves=60, maxiter=200))
    return probs


def batch_workload(a):
    """64 independent solve_k problems of uneven grid, eps^-1, omega and mode count in one
    omx_solve_k_batch call (launch groups per grid), against the same problems solved one
    after another with omx_solve_k."""
    import numpy as np
    import omx
    probs = batch_problems()
    nmodes = sum(p["nev"] for p in probs)
    omx.solve_k_batch(probs[:4], precision=a.precision)  # warm-up
    t0 = time.perf_counter()
    res = omx.solve_k_batch(probs, precision=a.precision)
    elapsed = time.perf_counter() - t0
    conv = sum(p["nev"] for p, r in zip(probs, res) if r["status"] == 0 and r["info"]["converged"])
    # the same problems one at a time through the single-solve entry point
    solvers = {}
    seq_k = []
    t1 = time.perf_counter()
    for p in probs:
        key = (p["grid"].Nx, p["grid"].Ny)
        sv = solvers.setdefault(key, omx.HIPSolver(precision=a.precision))
        km, _ = omx.solve_k(p["omega"], p["epsinv"], p["grid"], sv, nev=p["nev"], k_tol=p["k_tol"],
                            eig_tol=p["eig_tol"], max_eigsolves=p["max_eigsolves"], maxiter=p["maxiter"],
                            fetch_evecs=False)
        seq_k.append(km)
    seq = time.perf_counter() - t1
    agree = max(float(np.max(np.abs(r["kmags"] / k - 1))) for r, k in zip(res, seq_k))
    # two problems against a cold FP64 solve at a tighter tolerance
    errs = []
    for j in range(2):
        p = probs[j]
        kr, _ = omx.solve_k(p["omega"], p["epsinv"], p["grid"], omx.HIPSolver(precision="fp64"), nev=p["nev"],
                            k_tol=1e-12, eig_tol=1e-10, fetch_evecs=False)
        errs.append(float(np.max(np.abs(res[j]["kmags"] ** 2 - kr ** 2) / kr ** 2)))
    out = {
        "metric": "converged modes/s, 64-instance mixed-grid batch (128^2..384^2, 4-16 modes)",
        "value": round(conv / elapsed, 3),
        "unit": "modes/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 1,
        "ms_per_step": round(1e3 * elapsed, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64" if a.precision == "mixed" else "fp64",
        "data": "synthetic",
        "config": {
            "workload": "omx_solve_k_batch over 64 seeded ridge problems (grids 128^2..384^2, omega U(0.5,1), "
                        "nev 4..16, k_tol 1e-10, eig_tol 1e-8): launch groups per grid on one stream",
            "modes_total": nmodes,
            "modes_converged": conv,
            "one_by_one_ms": round(1e3 * seq, 1),
            "one_by_one_modes_per_s": round(nmodes / seq, 3),
            "batch_speedup": round(seq / elapsed, 2),
            "max_rel_diff_k_vs_one_by_one": agree,
            "max_rel_err_k2_vs_cold_fp64_first_two": max(errs),
        },
    }
    print(json.dumps(out), flush=True)


def copy_workload(a):
    """A plain device copy of 256 MiB (torch's copy kernel): each dispatch reads and writes
    exactly 262144 KiB, which calibrates rocprofv3's FETCH_SIZE / WRITE_SIZE counters, and its
    rate is the copy bandwidth the box reaches (the practical ceiling of an HBM-bound pass)."""
    import torch
    n = 256 << 20
    src = torch.empty(n // 4, dtype=torch.float32, device="cuda").uniform_()
    dst = torch.empty_like(src)
    for _ in range(max(1, a.warmup)):
        dst.copy_(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        dst.copy_(src)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / a.steps
    gbs = 2 * n / dt / 1e9
    print(json.dumps({"metric": "device copy bandwidth (read + write bytes / s)", "value": round(gbs, 1),
                      "unit": "GB/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
                      "ms_per_step": round(1e3 * dt, 4), "higher_is_better": True, "scaling": "weak",
                      "vs_baseline": None, "dtype": "f32", "data": "synthetic",
                      "config": {"workload": "copy of 256 MiB (64 Mi float32) on the device", "bytes_per_step": 2 * n},
                      "frac_of_peak": round(gbs / HBM_PEAK_GBS, 4)}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--no-plain", action="store_true", help="skip the plain-algorithm comparison solve")
    ap.add_argument("--no-learned", action="store_true", help="skip the comparison solves from the other start")
    ap.add_argument("--start", default="auto", choices=["auto", "learned", "plain"],
                    help="start of the timed solves: the shipped warm-start network (auto: when it loads) or plain")
    ap.add_argument("--tune", default="off", choices=["off", "on"],
                    help="on: omx_autotune's plan and partitions for the grid (wisdom file) before timing")
    ap.add_argument("--workload", default="solve_k", choices=["solve_k", "eme", "periodic", "batch", "copy"],
                    help="solve_k: the 256x256 TFLN modes/s line (default); eme: the 512x256 EME slice chain; "
                         "periodic: solve_k on a 128x96x24 3D Bragg-grating cell; batch: the 64-instance "
                         "mixed-grid batch; copy: a plain 256 MiB device copy (the HBM counters' calibration "
                         "and the copy bandwidth the box reaches)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}): every solve here fits one MI355X and "
              "multi-GPU sharding is by independent instances, which this library does not launch.", file=sys.stderr)
        sys.exit(2)
    if a.workload == "eme":
        eme_workload(a)
        return
    if a.workload == "periodic":
        periodic_workload(a)
        return
    if a.workload == "batch":
        batch_workload(a)
        return
    if a.workload == "copy":
        copy_workload(a)
        return

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
    tuned = None
    if a.tune == "on":
        from omx import tuning
        tuned = tuning.autotune(solver, grid, ei, NEV)  # outside the timed region, as FFTW planning
    kw = dict(nev=NEV, k_tol=1e-10, eig_tol=1e-8, max_eigsolves=60, maxiter=200, fetch_evecs=False,
              return_info=True)
    # the timed solves start from the shipped warm-start network (the guess is inside the
    # timed region); without it, or with --start plain, from k_guess and random vectors
    st = None if a.start == "plain" else load_learned()
    if a.start == "learned" and st is None:
        raise RuntimeError("--start learned: the warm-start network could not be loaded")

    def step(sv):
        if st is not None:
            from omx import warmstart as WS
            return WS.solve_k_learned(OMEGA, ei, grid, sv, st, **kw)
        return omx.solve_k(OMEGA, ei, grid, sv, **kw)

    for _ in range(a.warmup):
        step(solver)
    # reset the live kernel statistics so that they cover the timed region only
    h = solver.handle(grid, ei)
    omx._lib.check(omx._lib.lib().omx_set_option(h.h, b"profile", 1.0), h.h)

    converged = 0
    infos = []
    errs = []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        km, _, info = step(solver)
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
    # the other start on the same kernels, outside the timed region
    other = None
    if not a.no_learned:
        other = (plain_start(a, OMEGA, ei, grid, kw, slope_ref, k_ref, value, infos) if st is not None else
                 learned_start(a, OMEGA, ei, grid, kw, slope_ref, k_ref, value, infos))
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
            # FP64 phase: Rayleigh-Ritz steps of the eigenpair refinement (LOBPCG iterations with refine=0)
            "fp64_rr_steps_per_step": float(np.mean([i["lobpcg_iters_fp64"] for i in infos])),
            "newton_steps_per_step": float(np.mean([i["n_newton"] for i in infos])),
            "matvecs_per_step": float(np.mean([i["n_matvec"] for i in infos])),
            "parallelism": "single GPU; bands batched as independent Newton instances",
            "autotuned": tuned,
            "start": "learned (omx/data/warmnet_tfln256.npz, guess inside the timed region)" if st is not None
                     else "plain (k_guess, random vectors)",
        },
        "plain": plain,
        ("plain_start" if st is not None else "learned_start"): other,
        "roofline": roofline,
        "cpu_baseline": None if a.no_cpu_baseline else cpu_baseline(),
    }
    print(json.dumps(out), flush=True)


def load_learned():
    """The shipped warm-start network (trained on 512 solver-converged rib draws), or None."""
    from omx import warmstart as WS
    path = os.path.join(WS.DATA_DIR, "warmnet_tfln256.npz")
    if not os.path.exists(path):
        return None
    return WS.LearnedStart.load(path)


def plain_start(a, omega, ei, grid, kw, slope_ref, k_ref, value, infos):
    """The same solve from the plain start (k_guess, random vectors) on a fresh handle,
    outside the headline's timed region."""
    import numpy as np
    import omx
    solver = omx.HIPSolver(precision=a.precision)
    for _ in range(max(1, a.warmup)):
        omx.solve_k(omega, ei, grid, solver, **kw)
    conv, pi = 0, []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        km, _, info = omx.solve_k(omega, ei, grid, solver, **kw)
        conv += int(np.sum(np.abs(slope_ref * (km - k_ref)) / omega ** 2 <= REL_TOL))
        pi.append(info)
    dt = time.perf_counter() - t0
    return {"modes_per_s": round(conv / dt, 3), "ms_per_step": round(1e3 * dt / a.steps, 2),
            "modes_converged": conv, "learned_vs_plain_start": round(value / (conv / dt), 3) if conv else None,
            "device_ms_per_step": round(float(np.mean([i["device_ms"] for i in pi])), 2),
            "device_ms_per_step_learned_start": round(float(np.mean([i["device_ms"] for i in infos])), 2),
            "lobpcg_iters_fp32_per_step": float(np.mean([i["lobpcg_iters_fp32"] for i in pi])),
            "lobpcg_iters_fp32_per_step_learned_start": float(np.mean([i["lobpcg_iters_fp32"] for i in infos])),
            "newton_steps_per_step": float(np.mean([i["n_newton"] for i in pi]))}


def learned_start(a, omega, ei, grid, kw, slope_ref, k_ref, value, infos):
    """The same solve started from the trained warm-start network, the guess inside the
    timed region, when the headline ran from the plain start.  Outside the headline's timed
    region."""
    import numpy as np
    import omx
    from omx import warmstart as WS
    st = load_learned()
    if st is None:
        return None
    solver = omx.HIPSolver(precision=a.precision)
    for _ in range(max(1, a.warmup)):
        WS.solve_k_learned(omega, ei, grid, solver, st, **kw)
    conv, li = 0, []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        km, _, info = WS.solve_k_learned(omega, ei, grid, solver, st, **kw)
        conv += int(np.sum(np.abs(slope_ref * (km - k_ref)) / omega ** 2 <= REL_TOL))
        li.append(info)
    dt = time.perf_counter() - t0
    return {"modes_per_s": round(conv / dt, 3), "ms_per_step": round(1e3 * dt / a.steps, 2),
            "modes_converged": conv, "vs_plain_start": round((conv / dt) / value, 3) if value else None,
            "device_ms_per_step": round(float(np.mean([i["device_ms"] for i in li])), 2),
            "device_ms_per_step_plain_start": round(float(np.mean([i["device_ms"] for i in infos])), 2),
            "lobpcg_iters_fp32_per_step": float(np.mean([i["lobpcg_iters_fp32"] for i in li])),
            "lobpcg_iters_fp32_per_step_plain_start": float(np.mean([i["lobpcg_iters_fp32"] for i in infos])),
            "newton_steps_per_step": float(np.mean([i["n_newton"] for i in li]))}


if __name__ == "__main__":
    main()
