"""Autotuner wisdom for the HIP backend.

``omx_autotune`` times, on one handle's grid, the kernels of an nev-mode ``solve_k``: the
operator path (register FFT, with its block shapes, or LDS Stockham), the Stockham FFT radix
variant and tile shape,
on 3D cells the yz-slab or five-pass form, and the Gram and block-update row partitions, and
makes the fastest the handle's options.  This module keeps those choices per (device, grid
shape, nev) in a JSON file, as FFTW keeps its wisdom (the reference plans with
``FFTW.PATIENT`` when it builds a ``ModeSolver``, lib/MaxwellEigenmodes/src/maxwell.jl:940-943),
so that later processes apply them without timing again.

    choice = omx.tuning.autotune(solver, grid, epsinv, nev)   # tunes once, then reuses
"""
import ctypes
import json
import os

import numpy as np

from . import _lib

KEYS = ("reg_ops", "fft_variant32", "tile_rows32", "tile_cols32", "gram_blocks", "combine_blocks", "slab3d",
        "reg_rows32", "reg_nt2")


def default_wisdom_path():
    return os.environ.get("OMX_WISDOM", os.path.join(os.path.expanduser("~"), ".cache", "omx", "wisdom.json"))


def _key(grid, nev, device):
    nz = getattr(grid, "Nz", 1)
    return f"dev{int(device)}:{grid.Nx}x{grid.Ny}x{nz}:nev{int(nev)}"


def load(path=None):
    path = path or default_wisdom_path()
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def save(wisdom, path=None):
    path = path or default_wisdom_path()
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(wisdom, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def apply(solver, grid, epsinv, choice):
    """Set a stored choice on the solver's handle for this grid."""
    h = solver.handle(grid, epsinv)
    L = _lib.lib()
    for k in KEYS:
        _lib.check(L.omx_set_option(h.h, k.encode(), float(choice[k])), h.h)


def tune(solver, grid, epsinv, nev):
    """Run ``omx_autotune`` on the solver's handle for this grid: returns the choice (the
    options in ``KEYS``) with the measured microseconds of the chosen operator application,
    Gram and block update (``op_us``, ``gram_us``, ``combine_us``)."""
    h = solver.handle(grid, epsinv)
    c = (ctypes.c_int * len(KEYS))()
    us = np.zeros(3)
    _lib.check(_lib.lib().omx_autotune(h.h, int(nev), c, _lib.dptr(us)), h.h)
    out = {k: int(v) for k, v in zip(KEYS, c)}
    out.update(op_us=float(us[0]), gram_us=float(us[1]), combine_us=float(us[2]))
    return out


def autotune(solver, grid, epsinv, nev, wisdom_path=None, retune=False):
    """The stored choice for (device, grid shape, nev) applied to the solver's handle, or,
    when there is none (or ``retune``), a fresh ``tune`` whose result is stored."""
    w = load(wisdom_path)
    key = _key(grid, nev, solver.device)
    if not retune and key in w and all(k in w[key] for k in KEYS):
        apply(solver, grid, epsinv, w[key])
        return dict(w[key], cached=True)
    choice = tune(solver, grid, epsinv, nev)
    w[key] = choice
    try:
        save(w, wisdom_path)
    except OSError:
        pass  # a read-only home: the choice still holds for this handle
    return dict(choice, cached=False)
