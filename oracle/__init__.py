"""FP64 CPU oracle for the plane-wave Maxwell eigensolver hot path.

TEST INFRASTRUCTURE ONLY.  This package is a NumPy restatement of the reference
(OptiMode.jl, ``lib/MaxwellEigenmodes``) algorithms that the HIP library replaces.  It
is the reference point of the tests: only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it.  The product path (``optimode.jl_amd``) never
imports, calls or falls back to anything in here.

Every function cites the reference ``file:line`` it restates.  How each one is pinned
(closed forms, independent dense assembly, brute force, finite differences) is stated
by its tests in ``tests/test_oracle*.py``; a function
without such a check says "parity unpinned" in its docstring.
"""
from .maxwell import *  # noqa: F401,F403
from .lobpcg import lobpcg, ortho, ortho_against  # noqa: F401
from .solve import solve_omega2, solve_k, solve_k_single  # noqa: F401
from .adjoint import (eig_adjt, epsinv_bar, grad_k_mag_m_n, solve_k_pullback,  # noqa: F401
                      sliceinv_3x3, sliceinv_3x3_pullback, group_index)
from .analysis import ng_gvd, gaussian_wg_eps_dispersive  # noqa: F401
from . import periodic  # noqa: F401
