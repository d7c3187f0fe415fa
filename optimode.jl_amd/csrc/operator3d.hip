// Plane-wave Helmholtz operator on 3D cells periodic along z (Grid{3}), on the LDS
// Stockham FFT.  The steps are those of the 2D operator with 3D transforms: plane wave
// (ix, iy, iz) has w = k_vec - G = (-g_x, -g_y, k - g_z) (maxwell.jl:491-505), so plane iz
// of a vector uses the 2D basis at k - g_z(iz).  A vector is nz consecutive planes of
// 2 nx ny values (flat index c + 2(ix + nx(iy + ny iz))).  Two forms over a batch of
// vectors, with the 3-component intermediate field in scratch:
//   yz-slab form (the 3 ny nz values of an x-column fit one CU's LDS), three kernels,
//   intermediate x-major [v][ix][iz][iy][c]:
//     rows (per plane): k-space map + FFT along x
//     slab (per x-column): FFT along y and z, eps^-1 (or eps_ave), inverse FFT along z and y
//     rows (per plane): inverse FFT along x + map back onto (m, n), or the MkH dot
//   five-pass form (any cell), intermediate [v][iz][iy][ix][c]:
//     rows, cols (FFT along y per plane), pencils (FFT along z, eps^-1, inverse FFT along z
//     per x-run of one iy), cols (inverse FFT along y), rows
#include <algorithm>

#include "omx_errors.hpp"
#include "operator.hpp"

namespace omx {

namespace {

constexpr int B3 = 256;

__device__ __forceinline__ bool selected3(const VecSel& sel, int inst, int col) {
  if (sel.col0 == nullptr) return true;
  const int c0 = sel.col0[inst];
  return col >= c0 && col < c0 + sel.ncol[inst];
}

// transform directions: M and the MkH forms go forward first, P backward first
template <int OP> struct Dirs3 { static constexpr int first = -1, second = +1; };
template <> struct Dirs3<OP_P> { static constexpr int first = +1, second = -1; };

// rows of plane iz: k-space map of (h1, h2) at k - g_z, FFT along x
template <typename T, int OP, int VPT>
__global__ __launch_bounds__(B3) void k3_rows1(OpPlan<T> plan, BView<T> X, cplx<T>* D, VecSel sel,
                                               const double* kinst) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
  const int nx = plan.nx, ny = plan.ny, nz = plan.nz, RB = plan.rows_per_block;
  cplx<T>* buf = tw + nx;
  const int vz = blockIdx.y;
  const int v = vz / nz, iz = vz - (vz / nz) * nz;
  const int inst = v / sel.mper, col = v - (v / sel.mper) * sel.mper;
  if (!selected3(sel, inst, col)) return;
  const int iy0 = blockIdx.x * RB;
  for (int t = threadIdx.x; t < nx; t += blockDim.x) tw[t] = plan.twx[t];
  const double gz = double(fft_index(iz, nz)) * plan.invdz;
  const T k = T(kinst[inst] - gz);
  const T wt = T(gz * plan.invdz);  // OP_MKHW: g_z / Dz
  const cplx<T>* Xv = X.base + inst * X.istride + col * X.ld + long(iz) * 2 * nx * ny;
  for (int e = threadIdx.x; e < RB * nx; e += blockDim.x) {
    const int r = e / nx;
    const int ix = e - r * nx;
    const int iy = iy0 + r;
    const long p = ix + long(nx) * iy;
    const cplx<T> h1 = Xv[2 * p], h2 = Xv[2 * p + 1];
    const Basis<T> b = basis_at<T>(ix, iy, nx, ny, plan.invdx, plan.invdy, k);
    cplx<T> d0, d1, d2;
    if (OP == OP_M) {
      d0 = (h2 * b.m0 - h1 * b.n0) * b.mag;
      d1 = (h2 * b.m1 - h1 * b.n1) * b.mag;
      d2 = (h2 * b.m2 - h1 * b.n2) * b.mag;
    } else if (OP == OP_P) {
      const T im = T(1) / b.mag;
      d0 = (h1 * b.n0 - h2 * b.m0) * im;
      d1 = (h1 * b.n1 - h2 * b.m1) * im;
      d2 = (h1 * b.n2 - h2 * b.m2) * im;
    } else {
      // zx_tc (maxwell.jl:220-227), weighted by g_z / Dz for OP_MKHW (period.jl:69-74)
      const T f = OP == OP_MKHW ? wt : T(1);
      d0 = (h1 * b.m1 + h2 * b.n1) * (-f);
      d1 = (h1 * b.m0 + h2 * b.n0) * f;
      d2 = cplx<T>{T(0), T(0)};
    }
    buf[(r * 3 + 0) * nx + ix] = d0;
    buf[(r * 3 + 1) * nx + ix] = d1;
    buf[(r * 3 + 2) * nx + ix] = d2;
  }
  __syncthreads();
  fft_lds<T, Dirs3<OP>::first, VPT>(buf, 3 * RB, nx, 1, plan.fx, tw);
  cplx<T>* Dv = D + long(vz) * 3 * nx * ny;
  for (int e = threadIdx.x; e < RB * nx * 3; e += blockDim.x) {
    const int c = e % 3;
    const int q = e / 3;
    const int r = q / nx;
    const int ix = q - r * nx;
    Dv[(long(iy0 + r) * nx + ix) * 3 + c] = buf[(r * 3 + c) * nx + ix];
  }
}

// columns of plane iz: FFT along y in direction DIR, in place
template <typename T, int DIR, int VPT>
__global__ __launch_bounds__(B3) void k3_cols(OpPlan<T> plan, cplx<T>* D, VecSel sel) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
  const int nx = plan.nx, ny = plan.ny, nz = plan.nz, TX = plan.cols_per_block;
  cplx<T>* buf = tw + ny;
  const int vz = blockIdx.y;
  const int v = vz / nz;
  const int inst = v / sel.mper, col = v - (v / sel.mper) * sel.mper;
  if (!selected3(sel, inst, col)) return;
  const int ix0 = blockIdx.x * TX;
  for (int t = threadIdx.x; t < ny; t += blockDim.x) tw[t] = plan.twy[t];
  cplx<T>* Dv = D + long(vz) * 3 * nx * ny;
  for (int e = threadIdx.x; e < ny * TX * 3; e += blockDim.x) {
    const int c = e % 3;
    const int q = e / 3;
    const int iy = q / TX;
    const int t = q - iy * TX;
    buf[(t * 3 + c) * ny + iy] = Dv[(long(iy) * nx + ix0 + t) * 3 + c];
  }
  __syncthreads();
  fft_lds<T, DIR, VPT>(buf, 3 * TX, ny, 1, plan.fy, tw);
  for (int e = threadIdx.x; e < ny * TX * 3; e += blockDim.x) {
    const int c = e % 3;
    const int q = e / 3;
    const int iy = q / TX;
    const int t = q - iy * TX;
    Dv[(long(iy) * nx + ix0 + t) * 3 + c] = buf[(t * 3 + c) * ny + iy];
  }
}

// z-pencils of TZ consecutive ix at one iy.  MODE 0 (operator): FFT along z, the per-voxel
// eps^-1 (eps_ave for OP_P), inverse FFT along z.  Field helpers: MODE 1 (FIELD_E) forward
// FFT and eps^-1 times `scale`; MODE 2 (FIELD_D) forward FFT times `scale`; MODE 3
// (FIELD_EINV) eps^-1 times `scale`, then the backward FFT.
template <typename T, int OP, int VPT, int MODE>
__global__ __launch_bounds__(B3) void k3_pencils(OpPlan<T> plan, cplx<T>* D, VecSel sel, T scale) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
  const int nx = plan.nx, ny = plan.ny, nz = plan.nz, TZ = plan.zcols_per_block;
  cplx<T>* buf = tw + nz;
  const int v = blockIdx.y;
  const int inst = v / sel.mper, col = v - (v / sel.mper) * sel.mper;
  if (!selected3(sel, inst, col)) return;
  const int nxb = nx / TZ;
  const int iy = blockIdx.x / nxb;
  const int ix0 = (blockIdx.x - iy * nxb) * TZ;
  for (int t = threadIdx.x; t < nz; t += blockDim.x) tw[t] = plan.twz[t];
  cplx<T>* Dv = D + long(v) * 3 * nx * ny * nz;
  for (int e = threadIdx.x; e < nz * TZ * 3; e += blockDim.x) {
    const int c = e % 3;
    const int q = e / 3;
    const int iz = q / TZ;
    const int t = q - iz * TZ;
    buf[(t * 3 + c) * nz + iz] = Dv[((long(iz) * ny + iy) * nx + ix0 + t) * 3 + c];
  }
  __syncthreads();
  if (MODE != 3) fft_lds<T, MODE == 0 ? Dirs3<OP>::first : -1, VPT>(buf, 3 * TZ, nz, 1, plan.fz, tw);
  for (int e = threadIdx.x; e < nz * TZ; e += blockDim.x) {
    const int iz = e / TZ;
    const int t = e - iz * TZ;
    const T* ep = plan.eps6 + ((long(iz) * ny + iy) * nx + ix0 + t) * 6;
    cplx<T>* b0 = buf + (t * 3 + 0) * nz + iz;
    cplx<T>* b1 = buf + (t * 3 + 1) * nz + iz;
    cplx<T>* b2 = buf + (t * 3 + 2) * nz + iz;
    const cplx<T> d0 = *b0, d1 = *b1, d2 = *b2;
    if (MODE == 2) {
      *b0 = d0 * scale;
      *b1 = d1 * scale;
      *b2 = d2 * scale;
    } else if (OP == OP_P) {
      const T ea = T(3) / (ep[0] + ep[1] + ep[2]);
      *b0 = d0 * ea;
      *b1 = d1 * ea;
      *b2 = d2 * ea;
    } else {
      const T exx = ep[0], eyy = ep[1], ezz = ep[2], exy = ep[3], exz = ep[4], eyz = ep[5];
      const T s = MODE == 0 ? T(1) : scale;
      *b0 = (d0 * exx + d1 * exy + d2 * exz) * s;
      *b1 = (d0 * exy + d1 * eyy + d2 * eyz) * s;
      *b2 = (d0 * exz + d1 * eyz + d2 * ezz) * s;
    }
  }
  __syncthreads();
  if (MODE == 0) fft_lds<T, Dirs3<OP>::second, VPT>(buf, 3 * TZ, nz, 1, plan.fz, tw);
  if (MODE == 3) fft_lds<T, +1, VPT>(buf, 3 * TZ, nz, 1, plan.fz, tw);
  for (int e = threadIdx.x; e < nz * TZ * 3; e += blockDim.x) {
    const int c = e % 3;
    const int q = e / 3;
    const int iz = q / TZ;
    const int t = q - iz * TZ;
    Dv[((long(iz) * ny + iy) * nx + ix0 + t) * 3 + c] = buf[(t * 3 + c) * nz + iz];
  }
}

// rows of plane iz: inverse FFT along x, map back onto (m, n) at k - g_z scaled by
// 1/(nx ny nz); the MkH forms dot the result with the input instead (partials per block:
// vector v, block iz * gridDim.x + blockIdx.x of dps)
template <typename T, int OP, int VPT>
__global__ __launch_bounds__(B3) void k3_rows3(OpPlan<T> plan, BView<T> Hin, BView<T> Y, cplx<T>* D, VecSel sel,
                                               const double* kinst, double* dot_partial, int dps) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
  const int nx = plan.nx, ny = plan.ny, nz = plan.nz, RB = plan.rows_per_block;
  cplx<T>* buf = tw + nx;
  const int vz = blockIdx.y;
  const int v = vz / nz, iz = vz - (vz / nz) * nz;
  const int inst = v / sel.mper, col = v - (v / sel.mper) * sel.mper;
  if (!selected3(sel, inst, col)) return;
  const int iy0 = blockIdx.x * RB;
  for (int t = threadIdx.x; t < nx; t += blockDim.x) tw[t] = plan.twx[t];
  const cplx<T>* Dv = D + long(vz) * 3 * nx * ny;
  for (int e = threadIdx.x; e < RB * nx * 3; e += blockDim.x) {
    const int c = e % 3;
    const int q = e / 3;
    const int r = q / nx;
    const int ix = q - r * nx;
    buf[(r * 3 + c) * nx + ix] = Dv[(long(iy0 + r) * nx + ix) * 3 + c];
  }
  __syncthreads();
  fft_lds<T, Dirs3<OP>::second, VPT>(buf, 3 * RB, nx, 1, plan.fx, tw);
  const double gz = double(fft_index(iz, nz)) * plan.invdz;
  const T k = T(kinst[inst] - gz);
  const T scale = T(1.0 / (double(nx) * double(ny) * double(nz)));
  constexpr bool DOT = OP == OP_MKH || OP == OP_MKHW;
  const long plane = long(iz) * 2 * nx * ny;
  const cplx<T>* Hv = Hin.base + inst * Hin.istride + col * Hin.ld + plane;
  cplx<T>* Yv = DOT ? nullptr : Y.base + inst * Y.istride + col * Y.ld + plane;
  double acc = 0.0;
  for (int e = threadIdx.x; e < RB * nx; e += blockDim.x) {
    const int r = e / nx;
    const int ix = e - r * nx;
    const int iy = iy0 + r;
    const Basis<T> b = basis_at<T>(ix, iy, nx, ny, plan.invdx, plan.invdy, k);
    const cplx<T> e0 = buf[(r * 3 + 0) * nx + ix];
    const cplx<T> e1 = buf[(r * 3 + 1) * nx + ix];
    const cplx<T> e2 = buf[(r * 3 + 2) * nx + ix];
    const cplx<T> en = e0 * b.n0 + e1 * b.n1 + e2 * b.n2;
    const cplx<T> em = e0 * b.m0 + e1 * b.m1 + e2 * b.m2;
    cplx<T> h1, h2;
    if (OP == OP_P) {
      const T s = scale / b.mag;
      h1 = en * s;
      h2 = em * (-s);
    } else {
      const T s = b.mag * scale;
      h1 = en * (-s);
      h2 = em * s;
    }
    const long p = ix + long(nx) * iy;
    if (DOT) {
      const cplx<T> a1 = Hv[2 * p], a2 = Hv[2 * p + 1];
      acc += double(a1.x) * double(h1.x) + double(a1.y) * double(h1.y) + double(a2.x) * double(h2.x) +
             double(a2.y) * double(h2.y);
    } else {
      Yv[2 * p] = h1;
      Yv[2 * p + 1] = h2;
    }
  }
  if (DOT) {
    __shared__ double red[B3];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = B3 / 2; s > 0; s >>= 1) {
      if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) dot_partial[long(v) * dps + long(iz) * gridDim.x + blockIdx.x] = -red[0];
  }
}

// rows of plane iz of the field helpers' intermediate: backward FFT along x, in place
template <typename T, int VPT>
__global__ __launch_bounds__(B3) void k3_rows_inv(OpPlan<T> plan, cplx<T>* D, VecSel sel) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
  const int nx = plan.nx, ny = plan.ny, nz = plan.nz, RB = plan.rows_per_block;
  cplx<T>* buf = tw + nx;
  const int vz = blockIdx.y;
  const int v = vz / nz;
  const int inst = v / sel.mper, col = v - (v / sel.mper) * sel.mper;
  if (!selected3(sel, inst, col)) return;
  const int iy0 = blockIdx.x * RB;
  for (int t = threadIdx.x; t < nx; t += blockDim.x) tw[t] = plan.twx[t];
  cplx<T>* Dv = D + long(vz) * 3 * nx * ny;
  for (int e = threadIdx.x; e < RB * nx * 3; e += blockDim.x) {
    const int c = e % 3;
    const int q = e / 3;
    const int r = q / nx;
    const int ix = q - r * nx;
    buf[(r * 3 + c) * nx + ix] = Dv[(long(iy0 + r) * nx + ix) * 3 + c];
  }
  __syncthreads();
  fft_lds<T, +1, VPT>(buf, 3 * RB, nx, 1, plan.fx, tw);
  for (int e = threadIdx.x; e < RB * nx * 3; e += blockDim.x) {
    const int c = e % 3;
    const int q = e / 3;
    const int r = q / nx;
    const int ix = q - r * nx;
    Dv[(long(iy0 + r) * nx + ix) * 3 + c] = buf[(r * 3 + c) * nx + ix];
  }
}

// ---------------------------------------------------------------------------------------
// yz-slab path (the cell's ny x nz planes of an x-column fit in LDS): three kernels, the
// intermediate field x-major [v][ix][iz][iy][c] so that a slab is one contiguous run.
// ---------------------------------------------------------------------------------------

// rows of plane iz: k-space map at k - g_z + FFT along x (row buffer padded to nx + 1 so
// that the transposed store reads LDS without bank conflicts), stored x-major: for each ix
// a run of RB rows x 3 components
template <typename T, int OP, int VPT>
__global__ __launch_bounds__(B3) void k3s_rows1(OpPlan<T> plan, BView<T> X, cplx<T>* D, VecSel sel,
                                                const double* kinst) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
  const int nx = plan.nx, ny = plan.ny, nz = plan.nz, RB = plan.rows_per_block, sx = nx + 1;
  cplx<T>* buf = tw + nx;
  const int vz = blockIdx.y;
  const int v = vz / nz, iz = vz - (vz / nz) * nz;
  const int inst = v / sel.mper, col = v - (v / sel.mper) * sel.mper;
  if (!selected3(sel, inst, col)) return;
  const int iy0 = blockIdx.x * RB;
  for (int t = threadIdx.x; t < nx; t += blockDim.x) tw[t] = plan.twx[t];
  const double gz = double(fft_index(iz, nz)) * plan.invdz;
  const T k = T(kinst[inst] - gz);
  const T wt = T(gz * plan.invdz);
  const cplx<T>* Xv = X.base + inst * X.istride + col * X.ld + long(iz) * 2 * nx * ny;
  for (int e = threadIdx.x; e < RB * nx; e += blockDim.x) {
    const int r = e / nx;
    const int ix = e - r * nx;
    const int iy = iy0 + r;
    const long p = ix + long(nx) * iy;
    const cplx<T> h1 = Xv[2 * p], h2 = Xv[2 * p + 1];
    const Basis<T> b = basis_at<T>(ix, iy, nx, ny, plan.invdx, plan.invdy, k);
    cplx<T> d0, d1, d2;
    if (OP == OP_M) {
      d0 = (h2 * b.m0 - h1 * b.n0) * b.mag;
      d1 = (h2 * b.m1 - h1 * b.n1) * b.mag;
      d2 = (h2 * b.m2 - h1 * b.n2) * b.mag;
    } else if (OP == OP_P) {
      const T im = T(1) / b.mag;
      d0 = (h1 * b.n0 - h2 * b.m0) * im;
      d1 = (h1 * b.n1 - h2 * b.m1) * im;
      d2 = (h1 * b.n2 - h2 * b.m2) * im;
    } else {
      const T f = OP == OP_MKHW ? wt : T(1);
      d0 = (h1 * b.m1 + h2 * b.n1) * (-f);
      d1 = (h1 * b.m0 + h2 * b.n0) * f;
      d2 = cplx<T>{T(0), T(0)};
    }
    buf[(r * 3 + 0) * sx + ix] = d0;
    buf[(r * 3 + 1) * sx + ix] = d1;
    buf[(r * 3 + 2) * sx + ix] = d2;
  }
  __syncthreads();
  fft_lds<T, Dirs3<OP>::first, VPT>(buf, 3 * RB, sx, 1, plan.fx, tw);
  cplx<T>* Dv = D + long(v) * 3 * nx * ny * nz;
  const int run = 3 * RB;
  for (int e = threadIdx.x; e < RB * nx * 3; e += blockDim.x) {
    const int ix = e / run;
    const int q = e - ix * run;
    const int r = q / 3, c = q - (q / 3) * 3;
    Dv[((long(ix) * nz + iz) * ny + iy0 + r) * 3 + c] = buf[(r * 3 + c) * sx + ix];
  }
}

// one block (SLAB_THREADS threads) per TS x-columns: the slab [t][c][iz][iy] in LDS; FFT
// along y, FFT along z, eps^-1 (eps_ave for OP_P) per voxel, inverse FFT along z, inverse
// FFT along y.  The host admits the slab path only when every FFT call of the slab meets
// the values-per-thread bound SLAB_VPT at this block size.
template <typename T, int OP>
__global__ __launch_bounds__(SLAB_THREADS) void k3s_slab(OpPlan<T> plan, cplx<T>* D, VecSel sel) {
  constexpr int VPT = SLAB_VPT;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<T>* twy = reinterpret_cast<cplx<T>*>(smem_raw);
  const int nx = plan.nx, ny = plan.ny, nz = plan.nz, TS = plan.slab_cols;
  cplx<T>* twz = twy + ny;
  cplx<T>* buf = twz + nz;
  const int v = blockIdx.y;
  const int inst = v / sel.mper, col = v - (v / sel.mper) * sel.mper;
  if (!selected3(sel, inst, col)) return;
  const int ix0 = blockIdx.x * TS;
  for (int t = threadIdx.x; t < ny; t += blockDim.x) twy[t] = plan.twy[t];
  for (int t = threadIdx.x; t < nz; t += blockDim.x) twz[t] = plan.twz[t];
  const int nyz = ny * nz;
  cplx<T>* Sv = D + long(v) * 3 * nx * ny * nz + long(ix0) * 3 * nyz;
  const int n = TS * nyz * 3;
  for (int e = threadIdx.x; e < n; e += blockDim.x) {
    const int c = e % 3;
    const int q = e / 3;
    const int t = q / nyz;
    const int w = q - t * nyz;  // iz * ny + iy
    buf[(t * 3 + c) * nyz + w] = Sv[e];
  }
  __syncthreads();
  fft_lds<T, Dirs3<OP>::first, VPT>(buf, 3 * TS * nz, ny, 1, plan.fy, twy);
  fft_lds2<T, Dirs3<OP>::first, VPT>(buf, 3 * TS * ny, ny, nyz, 1, ny, plan.fz, twz);
  for (int e = threadIdx.x; e < TS * nyz; e += blockDim.x) {
    const int t = e / nyz;
    const int w = e - t * nyz;
    const int iz = w / ny, iy = w - (w / ny) * ny;
    const T* ep = plan.eps6c + 6 * (iy + long(ny) * (ix0 + t + long(nx) * iz));
    cplx<T>* b0 = buf + (t * 3 + 0) * nyz + w;
    cplx<T>* b1 = buf + (t * 3 + 1) * nyz + w;
    cplx<T>* b2 = buf + (t * 3 + 2) * nyz + w;
    const cplx<T> d0 = *b0, d1 = *b1, d2 = *b2;
    if (OP == OP_P) {
      const T ea = T(3) / (ep[0] + ep[1] + ep[2]);
      *b0 = d0 * ea;
      *b1 = d1 * ea;
      *b2 = d2 * ea;
    } else {
      const T exx = ep[0], eyy = ep[1], ezz = ep[2], exy = ep[3], exz = ep[4], eyz = ep[5];
      *b0 = d0 * exx + d1 * exy + d2 * exz;
      *b1 = d0 * exy + d1 * eyy + d2 * eyz;
      *b2 = d0 * exz + d1 * eyz + d2 * ezz;
    }
  }
  __syncthreads();
  fft_lds2<T, Dirs3<OP>::second, VPT>(buf, 3 * TS * ny, ny, nyz, 1, ny, plan.fz, twz);
  fft_lds<T, Dirs3<OP>::second, VPT>(buf, 3 * TS * nz, ny, 1, plan.fy, twy);
  for (int e = threadIdx.x; e < n; e += blockDim.x) {
    const int c = e % 3;
    const int q = e / 3;
    const int t = q / nyz;
    const int w = q - t * nyz;
    Sv[e] = buf[(t * 3 + c) * nyz + w];
  }
}

// rows of plane iz from the x-major intermediate: inverse FFT along x, map back at k - g_z
// (or the MkH dot, partials as in k3_rows3)
template <typename T, int OP, int VPT>
__global__ __launch_bounds__(B3) void k3s_rows3(OpPlan<T> plan, BView<T> Hin, BView<T> Y, cplx<T>* D, VecSel sel,
                                                const double* kinst, double* dot_partial, int dps) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
  const int nx = plan.nx, ny = plan.ny, nz = plan.nz, RB = plan.rows_per_block, sx = nx + 1;
  cplx<T>* buf = tw + nx;
  const int vz = blockIdx.y;
  const int v = vz / nz, iz = vz - (vz / nz) * nz;
  const int inst = v / sel.mper, col = v - (v / sel.mper) * sel.mper;
  if (!selected3(sel, inst, col)) return;
  const int iy0 = blockIdx.x * RB;
  for (int t = threadIdx.x; t < nx; t += blockDim.x) tw[t] = plan.twx[t];
  const cplx<T>* Dv = D + long(v) * 3 * nx * ny * nz;
  const int run = 3 * RB;
  for (int e = threadIdx.x; e < RB * nx * 3; e += blockDim.x) {
    const int ix = e / run;
    const int q = e - ix * run;
    const int r = q / 3, c = q - (q / 3) * 3;
    buf[(r * 3 + c) * sx + ix] = Dv[((long(ix) * nz + iz) * ny + iy0 + r) * 3 + c];
  }
  __syncthreads();
  fft_lds<T, Dirs3<OP>::second, VPT>(buf, 3 * RB, sx, 1, plan.fx, tw);
  const double gz = double(fft_index(iz, nz)) * plan.invdz;
  const T k = T(kinst[inst] - gz);
  const T scale = T(1.0 / (double(nx) * double(ny) * double(nz)));
  constexpr bool DOT = OP == OP_MKH || OP == OP_MKHW;
  const long plane = long(iz) * 2 * nx * ny;
  const cplx<T>* Hv = Hin.base + inst * Hin.istride + col * Hin.ld + plane;
  cplx<T>* Yv = DOT ? nullptr : Y.base + inst * Y.istride + col * Y.ld + plane;
  double acc = 0.0;
  for (int e = threadIdx.x; e < RB * nx; e += blockDim.x) {
    const int r = e / nx;
    const int ix = e - r * nx;
    const int iy = iy0 + r;
    const Basis<T> b = basis_at<T>(ix, iy, nx, ny, plan.invdx, plan.invdy, k);
    const cplx<T> e0 = buf[(r * 3 + 0) * sx + ix];
    const cplx<T> e1 = buf[(r * 3 + 1) * sx + ix];
    const cplx<T> e2 = buf[(r * 3 + 2) * sx + ix];
    const cplx<T> en = e0 * b.n0 + e1 * b.n1 + e2 * b.n2;
    const cplx<T> em = e0 * b.m0 + e1 * b.m1 + e2 * b.m2;
    cplx<T> h1, h2;
    if (OP == OP_P) {
      const T s = scale / b.mag;
      h1 = en * s;
      h2 = em * (-s);
    } else {
      const T s = b.mag * scale;
      h1 = en * (-s);
      h2 = em * s;
    }
    const long p = ix + long(nx) * iy;
    if (DOT) {
      const cplx<T> a1 = Hv[2 * p], a2 = Hv[2 * p + 1];
      acc += double(a1.x) * double(h1.x) + double(a1.y) * double(h1.y) + double(a2.x) * double(h2.x) +
             double(a2.y) * double(h2.y);
    } else {
      Yv[2 * p] = h1;
      Yv[2 * p + 1] = h2;
    }
  }
  if (DOT) {
    __shared__ double red[B3];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = B3 / 2; s > 0; s >>= 1) {
      if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) dot_partial[long(v) * dps + long(iz) * gridDim.x + blockIdx.x] = -red[0];
  }
}

struct Geo3 {
  dim3 g1, g2, g3;
  size_t lds13, lds2, lds3;
};

template <typename T> Geo3 geometry3(const OpPlan<T>& plan, const VecSel& sel) {
  const int nvec = sel.ninst * sel.mper;
  const size_t cs = sizeof(cplx<T>);
  Geo3 g;
  g.g1 = dim3(plan.ny / plan.rows_per_block, nvec * plan.nz);
  g.g2 = dim3(plan.nx / plan.cols_per_block, nvec * plan.nz);
  g.g3 = dim3((plan.nx / plan.zcols_per_block) * plan.ny, nvec);
  g.lds13 = cs * (size_t(plan.nx) + size_t(plan.rows_per_block) * 3 * plan.nx);
  g.lds2 = cs * (size_t(plan.ny) + size_t(plan.cols_per_block) * 3 * plan.ny);
  g.lds3 = cs * (size_t(plan.nz) + size_t(plan.zcols_per_block) * 3 * plan.nz);
  return g;
}

template <typename T, int OP, int VPT>
void op3_vpt(const OpPlan<T>& plan, BView<T> X, BView<T> Y, cplx<T>* D, const VecSel& sel, const double* kinst,
             double* dot_partial, int dps, hipStream_t s) {
  const Geo3 g = geometry3(plan, sel);
  if (plan.slab_cols > 0) {
    const size_t cs = sizeof(cplx<T>);
    const size_t lr = cs * (size_t(plan.nx) + size_t(plan.rows_per_block) * 3 * (plan.nx + 1));
    const size_t ls = cs * (size_t(plan.ny) + plan.nz + size_t(plan.slab_cols) * 3 * plan.ny * plan.nz);
    const dim3 gs(plan.nx / plan.slab_cols, sel.ninst * sel.mper);
    hipLaunchKernelGGL((k3s_rows1<T, OP, VPT>), g.g1, dim3(B3), lr, s, plan, X, D, sel, kinst);
    hipLaunchKernelGGL((k3s_slab<T, OP>), gs, dim3(SLAB_THREADS), ls, s, plan, D, sel);
    hipLaunchKernelGGL((k3s_rows3<T, OP, VPT>), g.g1, dim3(B3), lr, s, plan, X, Y, D, sel, kinst, dot_partial,
                       dps);
    return;
  }
  hipLaunchKernelGGL((k3_rows1<T, OP, VPT>), g.g1, dim3(B3), g.lds13, s, plan, X, D, sel, kinst);
  hipLaunchKernelGGL((k3_cols<T, Dirs3<OP>::first, VPT>), g.g2, dim3(B3), g.lds2, s, plan, D, sel);
  hipLaunchKernelGGL((k3_pencils<T, OP, VPT, 0>), g.g3, dim3(B3), g.lds3, s, plan, D, sel, T(1));
  hipLaunchKernelGGL((k3_cols<T, Dirs3<OP>::second, VPT>), g.g2, dim3(B3), g.lds2, s, plan, D, sel);
  hipLaunchKernelGGL((k3_rows3<T, OP, VPT>), g.g1, dim3(B3), g.lds13, s, plan, X, Y, D, sel, kinst, dot_partial,
                     dps);
}

template <typename T, int OP>
void op3(const OpPlan<T>& plan, BView<T> X, BView<T> Y, cplx<T>* D, const VecSel& sel, const double* kinst,
         double* dot_partial, int dps, hipStream_t s) {
  switch (plan.vpt) {
    case 8: op3_vpt<T, OP, 8>(plan, X, Y, D, sel, kinst, dot_partial, dps, s); break;
    case 16: op3_vpt<T, OP, 16>(plan, X, Y, D, sel, kinst, dot_partial, dps, s); break;
    case 32: op3_vpt<T, OP, 32>(plan, X, Y, D, sel, kinst, dot_partial, dps, s); break;
    default: throw Error(OMX_ERR_INTERNAL, "operator plan has an unsupported values-per-thread bound");
  }
}

template <typename T, int VPT>
void fields3_vpt(FieldOp fop, const OpPlan<T>& plan, BView<T> X, cplx<T>* D, const VecSel& sel,
                 const double* kinst, T scale, hipStream_t s) {
  const Geo3 g = geometry3(plan, sel);
  if (fop == FIELD_EINV) {
    hipLaunchKernelGGL((k3_pencils<T, OP_M, VPT, 3>), g.g3, dim3(B3), g.lds3, s, plan, D, sel, scale);
    hipLaunchKernelGGL((k3_cols<T, +1, VPT>), g.g2, dim3(B3), g.lds2, s, plan, D, sel);
    hipLaunchKernelGGL((k3_rows_inv<T, VPT>), g.g1, dim3(B3), g.lds13, s, plan, D, sel);
    return;
  }
  hipLaunchKernelGGL((k3_rows1<T, OP_M, VPT>), g.g1, dim3(B3), g.lds13, s, plan, X, D, sel, kinst);
  hipLaunchKernelGGL((k3_cols<T, -1, VPT>), g.g2, dim3(B3), g.lds2, s, plan, D, sel);
  if (fop == FIELD_E)
    hipLaunchKernelGGL((k3_pencils<T, OP_M, VPT, 1>), g.g3, dim3(B3), g.lds3, s, plan, D, sel, scale);
  else
    hipLaunchKernelGGL((k3_pencils<T, OP_M, VPT, 2>), g.g3, dim3(B3), g.lds3, s, plan, D, sel, scale);
}

}  // namespace

template <typename T>
void launch_operator3d(OpKind op, const OpPlan<T>& plan, BView<T> X, BView<T> Y, cplx<T>* D, const VecSel& sel,
                       const double* kinst, double* dot_partial, int dps, hipStream_t stream) {
  switch (op) {
    case OP_M: op3<T, OP_M>(plan, X, Y, D, sel, kinst, dot_partial, dps, stream); break;
    case OP_P: op3<T, OP_P>(plan, X, Y, D, sel, kinst, dot_partial, dps, stream); break;
    case OP_MKH: op3<T, OP_MKH>(plan, X, Y, D, sel, kinst, dot_partial, dps, stream); break;
    case OP_MKHW: op3<T, OP_MKHW>(plan, X, Y, D, sel, kinst, dot_partial, dps, stream); break;
  }
  OMX_HIP_CHECK(hipGetLastError());
}

// Field helpers on 3D cells, layout [v][iz][iy][ix][c]: FIELD_E (D = eps^-1 . fft3(kx_tc(H))
// * scale), FIELD_D (D = fft3(kx_tc(H)) * scale) and FIELD_EINV (D <- bfft3(eps^-1 . D) * scale)
template <typename T>
void launch_fields3d(FieldOp fop, const OpPlan<T>& plan, BView<T> X, cplx<T>* D, const VecSel& sel,
                     const double* kinst, double scale, hipStream_t stream) {
  if (fop != FIELD_E && fop != FIELD_D && fop != FIELD_EINV)
    throw Error(OMX_ERR_UNSUPPORTED, "this field helper is built for 2D cells only");
  switch (plan.vpt) {
    case 8: fields3_vpt<T, 8>(fop, plan, X, D, sel, kinst, T(scale), stream); break;
    case 16: fields3_vpt<T, 16>(fop, plan, X, D, sel, kinst, T(scale), stream); break;
    case 32: fields3_vpt<T, 32>(fop, plan, X, D, sel, kinst, T(scale), stream); break;
    default: throw Error(OMX_ERR_INTERNAL, "operator plan has an unsupported values-per-thread bound");
  }
  OMX_HIP_CHECK(hipGetLastError());
}

template void launch_operator3d<float>(OpKind, const OpPlan<float>&, BView<float>, BView<float>, cplx<float>*,
                                       const VecSel&, const double*, double*, int, hipStream_t);
template void launch_operator3d<double>(OpKind, const OpPlan<double>&, BView<double>, BView<double>,
                                        cplx<double>*, const VecSel&, const double*, double*, int, hipStream_t);
template void launch_fields3d<double>(FieldOp, const OpPlan<double>&, BView<double>, cplx<double>*, const VecSel&,
                                      const double*, double, hipStream_t);

}  // namespace omx
