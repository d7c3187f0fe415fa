// Tall-skinny block kernels for LOBPCG: complex Gram products and block updates.
//
// Blocks are batched over instances: column c of instance i of a part starts at
// base + i*istride + (off_i + c)*ld, with off_i / count_i read from device shape arrays,
// so the column windows (locked vectors, active R and P) change without host round trips.
#pragma once
#include "omx_common.hpp"

namespace omx {

struct ShapeRef {
  const int* off;  // device per-instance column offset (null: fixed_off)
  const int* cnt;  // device per-instance column count (null: fixed_cnt)
  int fixed_off;
  int fixed_cnt;
};

template <typename T> struct Part {
  const cplx<T>* base;
  long ld;
  long istride;
  ShapeRef s;
};

constexpr int MAX_PARTS = 3;

template <typename T> struct ColList {
  int n;
  Part<T> p[MAX_PARTS];
};

// Output view with a per-instance column offset.
template <typename T> struct OutView {
  cplx<T>* base;
  long ld;
  long istride;
  const int* off;  // may be null
  int fixed_off;
};

// G_i = A_i^H B_i on the matrix cores, written as per-chunk partial sums (FP64) to
// partial[(i*nchunk + chunk)*pm*pm + a*pm + b] for a < 16 ceil(nA_i/16), b < 16 ceil(nB_i/16)
// (other entries untouched).  pm in {16, 32, 48}; ca, cb (<= pm, <= 48) bound the column
// counts of A and B.  b_prefix_of_a: B is the leading columns of A (cb <= 16), whose values
// are then loaded once.  hermitian: A^H B is Hermitian (B = M A, M Hermitian); only the
// 16 x 16 tiles on and above the tile diagonal are computed (the others stay untouched).
template <typename T>
void launch_gram(const ColList<T>& A, const ColList<T>& B, int ninst, long K, int nchunk, int pm, c64* partial,
                 hipStream_t stream, int ca, int cb, bool b_prefix_of_a = false, bool hermitian = false);

// out_i[e] = sum_chunk partial[(i*nchunk + chunk)*esz + e], fixed order (deterministic);
// instances with cnt[i] == 0 are skipped (cnt: device, may be null).
void launch_reduce(const c64* partial, int ninst, int nchunk, int esz, c64* out, hipStream_t stream,
                   const int* cnt = nullptr);

enum CombineMode : int {
  CMB_PLAIN = 0,     // O1 = Y C
  CMB_TWO = 1,       // O1 = Y C, O2 = Y2 C
  CMB_RESID = 2,     // O1 = Y C, O2 = Y2 C, O3 = O2 - O1 diag(lam), norm partials of O3
  CMB_SUBTRACT = 3,  // O1 = O1 - Y C   (in place)
};

struct CombineCoef {
  const double* C;   // per instance: C[i*cstride + a*ldc*2 .. ] complex (re,im) row a, col q
  long cstride;      // in complex elements
  int ldc;           // complex elements between rows of C
  const int* qcnt;   // per-instance output column count (null: qfixed)
  int qfixed;
  const double* lam; // CMB_RESID: lam[i*lstride + q]
  long lstride;
};

// Row-parallel block update; nq <= qmax in {4, 8, 16}.  norm_partial (CMB_RESID):
// norm_partial[(i*nblk + blk)*qmax + q] = sum over the block's rows of |O3|^2.
template <typename T>
void launch_combine(CombineMode mode, const ColList<T>& Y, const ColList<T>& Y2, const CombineCoef& cf, int qmax,
                    OutView<T> O1, OutView<T> O2, OutView<T> O3, int ninst, long K, int nblk, double* norm_partial,
                    hipStream_t stream);

// Deterministic pseudo-random normal fill of `ncol` columns (complex, unit variance per part).
template <typename T>
void launch_randn(cplx<T>* base, long ld, long istride, int ninst, int ncol, long K, uint64_t seed,
                  hipStream_t stream);

// FP32 <-> FP64 copies of ncol columns per instance.
template <typename TS, typename TD>
void launch_convert(const cplx<TS>* src, long lds, long istrs, cplx<TD>* dst, long ldd, long istrd, int ninst,
                    int ncol, long K, hipStream_t stream);

// Copy columns [off_i, off_i + cnt_i) of every instance from src to dst, and with the same
// columns from src2 to dst2 in the same launch when src2 is given.
template <typename T>
void launch_copy_cols(const cplx<T>* src, cplx<T>* dst, long ld, long istride, int ninst, const int* off,
                      const int* cnt, long K, hipStream_t stream, const cplx<T>* src2 = nullptr,
                      cplx<T>* dst2 = nullptr);

}  // namespace omx
