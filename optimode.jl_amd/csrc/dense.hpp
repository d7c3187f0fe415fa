// Small dense FP64 kernels of the LOBPCG driver: one workgroup per instance, all data in
// LDS (block sizes up to 48 = 3 x 16 columns).
#pragma once
#include "omx_common.hpp"

namespace omx {

// Shape fields: per-instance ints stored structure-of-arrays, field f of instance i at
// shape[f*ninst + i].  Written by the dense kernels, read by the block/operator kernels.
enum ShapeField : int {
  F_XOFF = 0,   // first active X column this iteration (nlocked at iteration start)
  F_NACT,       // active X/R columns this iteration (M - F_XOFF), 0 when idle
  F_ROFF,       // first column of the active R / AR window
  F_NP,         // P columns in Y this iteration
  F_NLOCK,      // nlocked after this iteration's locking
  F_NNEW,       // M - F_NLOCK (next iteration's active count), 0 when idle
  F_NEWLY,      // newly locked this iteration (offset of the next R window)
  F_NCOPY,      // X/AX columns to mirror into the other buffer (newly locked, or all on convergence)
  F_M,          // block size (0 when the instance is idle)
  F_ONE,        // constant 1 while active (0 when idle) -- single-column views
  F_EIG,        // target column (Newton band)
  F_NPN,        // P columns for the ortho of R this iteration (0 at the first iteration)
  F_NFIELDS
};

struct LState {
  int M;             // block size
  int ncheck;        // lock this many leading vectors to converge (n_conv_check)
  int niter;         // LOBPCG iteration counter (0 = fake first iteration)
  int nlocked;
  int prev_nlocked;
  int nact;          // M - prev_nlocked
  int newly_locked;
  int np;            // P columns in the current Y
  int converged;     // 1 once nlocked >= ncheck
  int active;        // 0 when the instance takes no part in the current solve
  int status;        // 0 ok, else an OMX_ERR_* code
  int finished;      // converged or stopped at maxiter; idle afterwards
  double tol;        // locking threshold on residual norms
  double eps;        // machine epsilon of the vector storage type
  double ortho_tol;  // 2 eps
};

struct NewtonState {
  double k;          // current k
  double omega2;     // target
  double k_tol;      // |Δω²| threshold (Roots.jl atol semantics)
  double f;          // last Δω²
  double slope;      // last ∂ω²/∂k
  double lam;        // last eigenvalue of the band
  int band;          // eigind (0-based)
  int nsolves;
  int max_solves;
  int done;          // 1 converged, 2 failed (iteration cap)
};

struct DenseLaunch {
  int ninst;
  int pm;            // leading dimension of Gram / coefficient matrices (16, 32, 48)
  int nchunk;        // Gram partial chunks per instance
  int nblk;          // combine blocks per instance (residual-norm partials)
  int qmax;          // combine column bound
  int maxiter;
  double jac_stop;   // Rayleigh-Ritz Jacobi: stop after a sweep whose rotations all have |sin| <= jac_stop
};

// Rayleigh-Ritz: G (py x py) from gram partials -> cX (py x nact), lam (nact).
// diag_mode: the fake first iteration (cX = I, lam = Re diag G).
void launch_dense_rr(const DenseLaunch& dl, LState* st, int* shape, const c64* gram_partial, double* Cx,
                     double* lam, int diag_mode, hipStream_t stream);

// Residual norms -> locking -> cP = (cX - e)[:, xn] orthogonalised against cX (DFTK).
void launch_dense_lock(const DenseLaunch& dl, LState* st, int* shape, const double* norm_partial,
                       const double* Cx, double* Cp, double* resid_hist, hipStream_t stream);

// Upper Cholesky of the Gram O = R^H R (n = count of list R) with DFTK's shifting; writes
// C = U^{-1} (row-major, ld pm) for R <- R U^{-1}.  cnt_field selects the column count.
void launch_dense_chol(const DenseLaunch& dl, LState* st, const int* shape, int cnt_field, const c64* gram_partial,
                       double* C, hipStream_t stream);

// R <- (R - Z Z^H R) U^{-1} coefficients from the Gram of [R Z]^H R, with R = F_NNEW
// columns and Z = X (F_M columns) + P (F_NPN); C is (nr + nz) x nr, rows in [R Z] order.
void launch_dense_orth(const DenseLaunch& dl, LState* st, const int* shape, const c64* gram_partial, int pmg,
                       double* C, hipStream_t stream);

// End of iteration: roll next-iteration shape fields into place.
void launch_advance(const DenseLaunch& dl, LState* st, int* shape, hipStream_t stream);

// Start of a solve: shapes for the fake first iteration of every active instance.
void launch_lobpcg_init(const DenseLaunch& dl, LState* st, int* shape, const int* active, const int* band,
                        hipStream_t stream);

// Newton step: f = lam - omega^2 from the band's Rayleigh quotient (gram partials of
// x^H A x), slope = 2 HMkH (dot partials), k update, done flags.
void launch_newton(int ninst, int mper, NewtonState* ns, LState* st, const c64* rq_partial, int rq_nchunk, int rq_pm,
                   const double* hmkh_partial, int hmkh_nblk, double* kinst, hipStream_t stream);

}  // namespace omx
