// Kernels of the FP64 eigenpair refinement (Rayleigh-Ritz with one correction vector per
// instance): block Gram products and block updates over [block | extra column], and the
// band column's residual.
#pragma once
#include "solver.hpp"

namespace omx {

constexpr int RF_MAXC = 17;                 // block columns (<= 16) + one correction
constexpr int RF_PAIRS = RF_MAXC * RF_MAXC;

// Columns of instance i: block columns c < n1[i] at blk + i*istride + c*ld, then (when
// extra != nullptr) one extra column at extra + i*ld.  Instances with n1[i] == 0 are idle.
struct RCols {
  const c64* blk;
  long ld, istride;
  const c64* extra;
};

// G_i[a + RF_MAXC b] = sum_rows conj(U_a) V_b for a < nu_i = n1[i] + (U.extra != 0) and
// b < nv_i; V's block part has n1[i] columns when vmode = 0 and none when vmode = 1 (V is
// then its extra column only).  `part` holds nblk * ninst * RF_PAIRS partials.  With B (and
// part_uu, as large as part), B_i = U^H U is formed in the same pass over the rows.
void launch_bgram(const RCols& U, const RCols& V, int vmode, const int* n1, int ninst, long K, int nblk, c64* part,
                  c64* G, hipStream_t s, c64* part_uu = nullptr, c64* B = nullptr);

// out_j = sum_l in_l C_i[l + RF_MAXC j] with in = U's columns (n1[i] block columns plus the
// extra column when present); outputs j < n1[i] (omode = 0: a block laid out like U's, at
// out + i*ostride + j*ld) or the single output j = 0 (omode = 1: out + i*ld).  `out` may
// be U's own block or extra column: a thread reads all inputs of its row before it writes.
void launch_bcombine(const RCols& U, const c64* C, const int* n1, int omode, c64* out, long ostride, int ninst,
                     long K, hipStream_t s);

// x_i = X column col[i], r_i = AX column col[i] - lam[i] x_i (compact, stride ld), and
// partials of |r_i|^2 (nblk per instance) for the instances with n1[i] > 0.
void launch_band_resid(const c64* X, const c64* AX, long ld, long istride, const int* col, const double* lam,
                       const int* n1, int ninst, long K, int nblk, c64* xb, c64* rb, double* part, hipStream_t s);

// x_i *= scale[i] (compact vectors)
void launch_scale_cols(c64* x, long ld, const double* scale, const int* n1, int ninst, long K, hipStream_t s);

}  // namespace omx
