// Fused plane-wave Helmholtz operator family (host-side launch interface).
//
//   M:   H -> kx_ct( bfft( eps^-1 . fft( kx_tc(H) ) ) ) / N            (maxwell.jl:471-481)
//   P:   H -> kxinv_ct( fft( eps_ave . bfft( kxinv_tc(H) ) ) ) / N       (maxwell.jl:459-469)
//   MkH: H -> kx_ct( ifft( eps^-1 . fft( zx_tc(H) ) ) )                  (maxwell.jl:267-275)
//
// Every application is three kernels over a batch of vectors:
//   pass 1 (rows):    pointwise curl in k-space + 1D FFT along x, all 3 components
//   pass 2 (columns): 1D FFT along y, 3x3 eps^-1 (or eps_ave) per pixel, inverse FFT along y
//   pass 3 (rows):    inverse 1D FFT along x + curl back onto the (m, n) basis
// with the intermediate 3-component field in a scratch buffer (L2/MALL-resident for the
// grid sizes of interest).  The basis (m, n, |k+G|) is recomputed on the fly from k and
// the integer reciprocal-lattice indices, so no per-k basis array is ever stored.
//
// Two implementations share this interface.  Power-of-two nx, ny (64..512) use
// register-resident FFTs (fft_reg.hpp: a group of L/8 lanes per 1D sequence, LDS
// exchanges between radix passes) with the intermediate field column-major per component,
// [v][c][ix][iy], so that the column pass reads and writes whole cache lines; eps^-1 is
// kept column-major as well.  Other sizes use the LDS Stockham FFT (fft_lds.hpp) with the
// intermediate field as [v][iy][ix][c].
#pragma once
#include "omx_common.hpp"
#include "fft_lds.hpp"
#include "fft_reg.hpp"

namespace omx {

// OP_MKHW (3D cells only): MkH with the zx_tc field weighted plane wave by plane wave by
// g_z / Dz, i.e. <H| dM/dDz |H> with Dz the period along z (grads/period.jl:69-74, :88-90)
enum OpKind : int { OP_M = 0, OP_P = 1, OP_MKH = 2, OP_MKHW = 3 };

template <typename T> struct Basis {
  T mag, m0, m1, m2, n0, n1, n2;
};

// Transverse basis at w = (0,0,k) - G: mag = |w|, n = normalize(z x w) (y at the gimbal
// point), m = normalize(n x w)   (maxwell.jl:646-660).
template <typename T>
__device__ __forceinline__ Basis<T> basis_at(int ix, int iy, int nx, int ny, double invdx, double invdy, T k) {
  const T gx = T(double(fft_index(ix, nx)) * invdx);
  const T gy = T(double(fft_index(iy, ny)) * invdy);
  const T wx = -gx, wy = -gy, wz = k;
  Basis<T> b;
  b.mag = sqrt(wx * wx + wy * wy + wz * wz);
  const T rho2 = wx * wx + wy * wy;
  if (rho2 > T(0)) {
    const T ir = T(1) / sqrt(rho2);
    b.n0 = -wy * ir;
    b.n1 = wx * ir;
  } else {
    b.n0 = T(0);
    b.n1 = T(1);
  }
  b.n2 = T(0);
  const T cx = b.n1 * wz - b.n2 * wy;
  const T cy = b.n2 * wx - b.n0 * wz;
  const T cz = b.n0 * wy - b.n1 * wx;
  const T ic = T(1) / sqrt(cx * cx + cy * cy + cz * cz);
  b.m0 = cx * ic;
  b.m1 = cy * ic;
  b.m2 = cz * ic;
  return b;
}

// Batched block view: column c of instance i starts at base + i*istride + c*ld.
template <typename T> struct BView {
  cplx<T>* base;
  long ld;
  long istride;
};

// Per-launch description of the vectors to transform.  Vector v of the launch is
// column (v % mper) of instance (v / mper); it is skipped unless
// col0[i] <= column < col0[i] + ncol[i] (both arrays may be null: all columns).
struct VecSel {
  int ninst;
  int mper;
  const int* col0;   // device, per instance, may be null
  const int* ncol;   // device, per instance, may be null
};

template <typename T> struct OpPlan {
  int nx, ny;
  double invdx, invdy;            // 1/Dx, 1/Dy: G = (fft_index(ix)/Dx, fft_index(iy)/Dy)
  FFTFactors fx, fy;
  const cplx<T>* twx;             // device: exp(-2 pi i t/nx), t < nx
  const cplx<T>* twy;             // device: exp(-2 pi i t/ny)
  const T* eps6;                  // device: (ny, nx, 6) = xx, yy, zz, xy, xz, yz
  const T* eps6c;                 // device: (nx, ny, 6), column-major copy (register path)
  // several eps^-1 fields on one grid (problems batched into one launch, 2D cells): vector v
  // of instance i reads the field at eps6 / eps6c + eps_slot[i] * eps_stride; null: one field
  const int* eps_slot;            // device, per instance, may be null
  long eps_stride;                // 6 N
  int rows_per_block;             // Stockham path: pass 1/3
  int cols_per_block;             // Stockham path: pass 2
  int vpt;                        // Stockham path: values per thread bound (8, 16, 32)
  int reg;                        // 1: the operator uses the register-FFT path
  int reg_rb;                     // register path: rows per block of passes 1/3 (0: reg_rows<T>())
  int reg_nt2;                    // register path: threads per block of pass 2 (0: REG_COL_THREADS)
  int dot_blocks;                 // OP_MKH dot partials per vector (blocks of pass 3)
  // 3D cells periodic along z (nz > 1; 2D: nz = 1): plane iz has g_z = fft_index(iz)/Dz,
  // vectors are nz consecutive planes of 2 nx ny values, the intermediate field is
  // [v][iz][iy][ix][c], eps6 is (nz, ny, nx, 6)
  int nz;
  double invdz;
  FFTFactors fz;
  const cplx<T>* twz;             // device: exp(-2 pi i t/nz)
  int zcols_per_block;            // x-columns per block of the z pass
  // > 0: the yz-slab path (whole x-columns of ny x nz values per block in LDS, intermediate
  // x-major [v][ix][iz][iy][c]), x-columns per block; 0: five passes
  int slab_cols;
};

// eps^-1 of instance `inst` in the two device layouts
template <typename T> __device__ __forceinline__ const T* eps_rowm(const OpPlan<T>& p, int inst) {
  return p.eps_slot ? p.eps6 + long(p.eps_slot[inst]) * p.eps_stride : p.eps6;
}
template <typename T> __device__ __forceinline__ const T* eps_colm(const OpPlan<T>& p, int inst) {
  return p.eps_slot ? p.eps6c + long(p.eps_slot[inst]) * p.eps_stride : p.eps6c;
}

// Register path geometry: rows per block of passes 1 and 3 (the column-major intermediate
// is written / read in runs of that many iy: whole 128-byte lines), threads per block of
// pass 1 / 3 (a group of L/8 lanes per row) and of pass 2 (a group per column).
// Each precision has a second, halved row count (more, smaller blocks for launches of few
// vectors) and pass 2 a halved thread count; plan.reg_rb / plan.reg_nt2 pick them.
template <typename T> constexpr int reg_rows() { return sizeof(T) == 4 ? 16 : 8; }
template <typename T, int LX> constexpr int reg_row_threads() { return reg_rows<T>() * (LX / 8); }
constexpr int REG_COL_THREADS = 256;
template <typename T> inline int reg_rows_of(const OpPlan<T>& p) { return p.reg_rb > 0 ? p.reg_rb : reg_rows<T>(); }

// Launch one operator application on `stream`.
//   X: input block, Y: output block (OP_M/OP_P), D: scratch of 3*N complex per vector
//   (column v of the launch at D + v*3N), kinst: per-instance k (device, FP64).
//   For OP_MKH, Y is unused and dot_partial[v * dot_partial_stride + b], b < plan.dot_blocks,
//   receive partial sums of -Re<H_v, MkH H_v> (FP64, device).
template <typename T>
void launch_operator(OpKind op, const OpPlan<T>& plan, BView<T> X, BView<T> Y, cplx<T>* D,
                     const VecSel& sel, const double* kinst, double* dot_partial, int dot_partial_stride,
                     hipStream_t stream);

// Field helpers on 3-component fields in D (layout [v][iy][ix][c], 3N per vector):
//   FIELD_E: D = eps^-1 . fft2(kx_tc(H)) * scale   (E field, fields.jl:194-210)
//   FIELD_D: D = fft2(kx_tc(H)) * scale             (_H2d!, fields.jl:81-86)
//   FIELD_EINV: D <- bfft2(eps^-1 . D) * scale      (_d2ẽ!, fields.jl:88-93), X unused
//   FIELD_DK: D = fft2(d kx_tc/dk (H)) * scale = -fft2(zx_tc(H)) * scale
//   FIELD_DKK: D = fft2(d^2 kx_tc/dk^2 (H)) * scale = fft2(-H1 n |G|^2 / |k+G|^3) * scale
//   FIELD_H: D = fft2(tc(H)) * scale = fft2(H1 m + H2 n) * scale   (modes.jl:112)
//   FIELD_BFFT: D <- bfft2(D) * scale, X unused (2D cells)
// (the k-derivatives are taken at fixed transverse coefficients H, with the basis
// m(k), n, |k+G|(k) of basis_at)
enum FieldOp : int { FIELD_E = 0, FIELD_D = 1, FIELD_EINV = 2, FIELD_DK = 3, FIELD_DKK = 4, FIELD_H = 5, FIELD_BFFT = 6 };
template <typename T>
void launch_fields(FieldOp fop, const OpPlan<T>& plan, BView<T> X, cplx<T>* D, const VecSel& sel,
                   const double* kinst, double scale, hipStream_t stream);

// yz-slab kernel geometry of the 3D operator: threads per block and values-per-thread bound
// of its FFT calls
constexpr int SLAB_THREADS = 512;
constexpr int SLAB_VPT = 16;

// 3D cells (plan.nz > 1): the same operator and the FIELD_E helper with 3D transforms;
// launch_operator / launch_fields forward here.
template <typename T>
void launch_operator3d(OpKind op, const OpPlan<T>& plan, BView<T> X, BView<T> Y, cplx<T>* D, const VecSel& sel,
                       const double* kinst, double* dot_partial, int dot_partial_stride, hipStream_t stream);
template <typename T>
void launch_fields3d(FieldOp fop, const OpPlan<T>& plan, BView<T> X, cplx<T>* D, const VecSel& sel,
                     const double* kinst, double scale, hipStream_t stream);

// Host helper: factor L into radices {8,4,3,5,2} (no lone factor 2 unless L is 2*odd).
bool fft_factor(int L, FFTFactors& f);
// The same length in radix-4 passes (then 3s and 5s) when its power of two is even and at
// least 4; otherwise as fft_factor.
bool fft_factor_radix4(int L, FFTFactors& f);

}  // namespace omx
