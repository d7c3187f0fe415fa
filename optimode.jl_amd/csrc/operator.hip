// Fused plane-wave Helmholtz operator kernels for gfx950 (see operator.hpp).
#include "operator.hpp"
#include <algorithm>
#include "omx_errors.hpp"

namespace omx {

namespace {

constexpr int OP_BLOCK = 256;

__device__ __forceinline__ bool vec_selected(const VecSel& sel, int inst, int col) {
  if (sel.col0 == nullptr) return true;
  const int c0 = sel.col0[inst];
  return col >= c0 && col < c0 + sel.ncol[inst];
}

template <int OP> struct Dirs;
template <> struct Dirs<OP_M> { static constexpr int x1 = -1, y1 = -1, y2 = +1, x2 = +1; };
template <> struct Dirs<OP_MKH> { static constexpr int x1 = -1, y1 = -1, y2 = +1, x2 = +1; };
template <> struct Dirs<OP_P> { static constexpr int x1 = +1, y1 = +1, y2 = -1, x2 = -1; };
// pass-1 maps of the field helpers FIELD_DKK (second k-derivative of kx_tc) and FIELD_H
// (tc: transverse to Cartesian); forward only
constexpr int CURL_DKK = 3;
constexpr int MAP_TC = 4;
template <> struct Dirs<CURL_DKK> { static constexpr int x1 = -1, y1 = -1, y2 = +1, x2 = +1; };
template <> struct Dirs<MAP_TC> { static constexpr int x1 = -1, y1 = -1, y2 = +1, x2 = +1; };

// ---------------------------------------------------------------------------------------
// pass 1: k-space map (kx_tc / kxinv_tc / zx_tc / d2 kx_tc / dk2 / tc) + FFT along x
// ---------------------------------------------------------------------------------------
template <typename T, int OP, int VPT>
__global__ __launch_bounds__(OP_BLOCK) void k_pass1(OpPlan<T> plan, BView<T> X, cplx<T>* D, VecSel sel,
                                                    const double* kinst) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
  const int nx = plan.nx, ny = plan.ny, RB = plan.rows_per_block;
  cplx<T>* buf = tw + nx;
  const int v = blockIdx.y;
  const int inst = v / sel.mper, col = v - (v / sel.mper) * sel.mper;
  if (!vec_selected(sel, inst, col)) return;
  const int iy0 = blockIdx.x * RB;
  for (int t = threadIdx.x; t < nx; t += blockDim.x) tw[t] = plan.twx[t];
  const T k = T(kinst[inst]);
  const cplx<T>* Xv = X.base + inst * X.istride + col * X.ld;
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
    } else if (OP == OP_MKH) {
      d0 = T(-1) * (h1 * b.m1 + h2 * b.n1);
      d1 = h1 * b.m0 + h2 * b.n0;
      d2 = cplx<T>{T(0), T(0)};
    } else if (OP == MAP_TC) {
      d0 = h1 * b.m0 + h2 * b.n0;
      d1 = h1 * b.m1 + h2 * b.n1;
      d2 = h1 * b.m2 + h2 * b.n2;
    } else {
      // d^2/dk^2 of kx_tc at fixed (h1, h2): -h1 n rho^2 / |k+G|^3, rho = |G|
      const T gx = T(double(fft_index(ix, nx)) * plan.invdx), gy = T(double(fft_index(iy, ny)) * plan.invdy);
      const T f = -(gx * gx + gy * gy) / (b.mag * b.mag * b.mag);
      d0 = h1 * (b.n0 * f);
      d1 = h1 * (b.n1 * f);
      d2 = cplx<T>{T(0), T(0)};
    }
    buf[(r * 3 + 0) * nx + ix] = d0;
    buf[(r * 3 + 1) * nx + ix] = d1;
    buf[(r * 3 + 2) * nx + ix] = d2;
  }
  __syncthreads();
  fft_lds<T, Dirs<OP>::x1, VPT>(buf, 3 * RB, nx, 1, plan.fx, tw);
  cplx<T>* Dv = D + long(v) * 3 * nx * ny;
  for (int e = threadIdx.x; e < RB * nx * 3; e += blockDim.x) {
    const int c = e % 3;
    const int q = e / 3;
    const int r = q / nx;
    const int ix = q - r * nx;
    Dv[(long(iy0 + r) * nx + ix) * 3 + c] = buf[(r * 3 + c) * nx + ix];
  }
}

// ---------------------------------------------------------------------------------------
// pass 2: FFT along y, pointwise eps^-1 (or eps_ave), inverse FFT along y
// ---------------------------------------------------------------------------------------
template <typename T, int OP, int VPT>
__global__ __launch_bounds__(OP_BLOCK) void k_pass2(OpPlan<T> plan, cplx<T>* D, VecSel sel) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
  const int nx = plan.nx, ny = plan.ny, TX = plan.cols_per_block;
  cplx<T>* buf = tw + ny;
  const int v = blockIdx.y;
  const int inst = v / sel.mper, col = v - (v / sel.mper) * sel.mper;
  if (!vec_selected(sel, inst, col)) return;
  const int ix0 = blockIdx.x * TX;
  for (int t = threadIdx.x; t < ny; t += blockDim.x) tw[t] = plan.twy[t];
  cplx<T>* Dv = D + long(v) * 3 * nx * ny;
  for (int e = threadIdx.x; e < ny * TX * 3; e += blockDim.x) {
    const int c = e % 3;
    const int q = e / 3;
    const int iy = q / TX;
    const int t = q - iy * TX;
    buf[(t * 3 + c) * ny + iy] = Dv[(long(iy) * nx + ix0 + t) * 3 + c];
  }
  __syncthreads();
  fft_lds<T, Dirs<OP>::y1, VPT>(buf, 3 * TX, ny, 1, plan.fy, tw);
  for (int e = threadIdx.x; e < ny * TX; e += blockDim.x) {
    const int iy = e / TX;
    const int t = e - iy * TX;
    const T* ep = eps_rowm(plan, inst) + (long(iy) * nx + ix0 + t) * 6;
    const T exx = ep[0], eyy = ep[1], ezz = ep[2], exy = ep[3], exz = ep[4], eyz = ep[5];
    cplx<T>* b0 = buf + (t * 3 + 0) * ny + iy;
    cplx<T>* b1 = buf + (t * 3 + 1) * ny + iy;
    cplx<T>* b2 = buf + (t * 3 + 2) * ny + iy;
    const cplx<T> d0 = *b0, d1 = *b1, d2 = *b2;
    if (OP == OP_P) {
      const T ea = T(3) / (exx + eyy + ezz);
      *b0 = d0 * ea;
      *b1 = d1 * ea;
      *b2 = d2 * ea;
    } else {
      *b0 = d0 * exx + d1 * exy + d2 * exz;
      *b1 = d0 * exy + d1 * eyy + d2 * eyz;
      *b2 = d0 * exz + d1 * eyz + d2 * ezz;
    }
  }
  __syncthreads();
  fft_lds<T, Dirs<OP>::y2, VPT>(buf, 3 * TX, ny, 1, plan.fy, tw);
  for (int e = threadIdx.x; e < ny * TX * 3; e += blockDim.x) {
    const int c = e % 3;
    const int q = e / 3;
    const int iy = q / TX;
    const int t = q - iy * TX;
    Dv[(long(iy) * nx + ix0 + t) * 3 + c] = buf[(t * 3 + c) * ny + iy];
  }
}

// ---------------------------------------------------------------------------------------
// pass 3: inverse FFT along x + curl back (kx_ct / kxinv_ct), scaled by 1/N; for OP_MKH
// the result is dotted with the input instead of stored.
// ---------------------------------------------------------------------------------------
template <typename T, int OP, int VPT>
__global__ __launch_bounds__(OP_BLOCK) void k_pass3(OpPlan<T> plan, BView<T> Hin, BView<T> Y, cplx<T>* D,
                                                    VecSel sel, const double* kinst, double* dot_partial,
                                                    int dps) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
  const int nx = plan.nx, ny = plan.ny, RB = plan.rows_per_block;
  cplx<T>* buf = tw + nx;
  const int v = blockIdx.y;
  const int inst = v / sel.mper, col = v - (v / sel.mper) * sel.mper;
  if (!vec_selected(sel, inst, col)) return;
  const int iy0 = blockIdx.x * RB;
  for (int t = threadIdx.x; t < nx; t += blockDim.x) tw[t] = plan.twx[t];
  const cplx<T>* Dv = D + long(v) * 3 * nx * ny;
  for (int e = threadIdx.x; e < RB * nx * 3; e += blockDim.x) {
    const int c = e % 3;
    const int q = e / 3;
    const int r = q / nx;
    const int ix = q - r * nx;
    buf[(r * 3 + c) * nx + ix] = Dv[(long(iy0 + r) * nx + ix) * 3 + c];
  }
  __syncthreads();
  fft_lds<T, Dirs<OP>::x2, VPT>(buf, 3 * RB, nx, 1, plan.fx, tw);
  const T k = T(kinst[inst]);
  const T scale = T(1.0 / (double(nx) * double(ny)));
  double acc = 0.0;
  const cplx<T>* Hv = Hin.base + inst * Hin.istride + col * Hin.ld;
  cplx<T>* Yv = (OP == OP_MKH) ? nullptr : Y.base + inst * Y.istride + col * Y.ld;
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
    if (OP == OP_MKH) {
      const cplx<T> a1 = Hv[2 * p], a2 = Hv[2 * p + 1];
      acc += double(a1.x) * double(h1.x) + double(a1.y) * double(h1.y) + double(a2.x) * double(h2.x) +
             double(a2.y) * double(h2.y);
    } else {
      Yv[2 * p] = h1;
      Yv[2 * p + 1] = h2;
    }
  }
  if (OP == OP_MKH) {
    // block reduction of acc (FP64), deterministic order
    __shared__ double red[OP_BLOCK];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
      if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) dot_partial[long(v) * dps + blockIdx.x] = -red[0];
  }
}

// Forward half-pass along y for the field helpers: FFT along y, optionally eps^-1 . d,
// times `scale`, left in D as [v][iy][ix][c].  With pass 1 in front this is
// fields.jl:197 ``ε⁻¹_dot(fft(kx_tc(H)))`` (EPS) or ``_H2d!`` (fields.jl:81-86, no EPS).
template <typename T, int VPT, bool EPS>
__global__ __launch_bounds__(OP_BLOCK) void k_pass2_fwd(OpPlan<T> plan, cplx<T>* D, VecSel sel, T scale) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
  const int nx = plan.nx, ny = plan.ny, TX = plan.cols_per_block;
  cplx<T>* buf = tw + ny;
  const int v = blockIdx.y;
  const int inst = v / sel.mper, col = v - (v / sel.mper) * sel.mper;
  if (!vec_selected(sel, inst, col)) return;
  const int ix0 = blockIdx.x * TX;
  for (int t = threadIdx.x; t < ny; t += blockDim.x) tw[t] = plan.twy[t];
  cplx<T>* Dv = D + long(v) * 3 * nx * ny;
  for (int e = threadIdx.x; e < ny * TX * 3; e += blockDim.x) {
    const int c = e % 3;
    const int q = e / 3;
    const int iy = q / TX;
    const int t = q - iy * TX;
    buf[(t * 3 + c) * ny + iy] = Dv[(long(iy) * nx + ix0 + t) * 3 + c];
  }
  __syncthreads();
  fft_lds<T, -1, VPT>(buf, 3 * TX, ny, 1, plan.fy, tw);
  for (int e = threadIdx.x; e < ny * TX; e += blockDim.x) {
    const int iy = e / TX;
    const int t = e - iy * TX;
    const cplx<T> d0 = buf[(t * 3 + 0) * ny + iy], d1 = buf[(t * 3 + 1) * ny + iy], d2 = buf[(t * 3 + 2) * ny + iy];
    const long o = (long(iy) * nx + ix0 + t) * 3;
    if (EPS) {
      const T* ep = eps_rowm(plan, inst) + (long(iy) * nx + ix0 + t) * 6;
      const T exx = ep[0], eyy = ep[1], ezz = ep[2], exy = ep[3], exz = ep[4], eyz = ep[5];
      Dv[o + 0] = (d0 * exx + d1 * exy + d2 * exz) * scale;
      Dv[o + 1] = (d0 * exy + d1 * eyy + d2 * eyz) * scale;
      Dv[o + 2] = (d0 * exz + d1 * eyz + d2 * ezz) * scale;
    } else {
      Dv[o + 0] = d0 * scale;
      Dv[o + 1] = d1 * scale;
      Dv[o + 2] = d2 * scale;
    }
  }
}

// Inverse path of _d2ẽ! (fields.jl:88-93): e = bfft2(eps^-1 . d) * scale, in place in D.
// Column pass: eps^-1 . d and inverse FFT along y.
template <typename T, int VPT, bool EPS>
__global__ __launch_bounds__(OP_BLOCK) void k_pass2_inv(OpPlan<T> plan, cplx<T>* D, VecSel sel, T scale) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
  const int nx = plan.nx, ny = plan.ny, TX = plan.cols_per_block;
  cplx<T>* buf = tw + ny;
  const int v = blockIdx.y;
  const int inst = v / sel.mper, col = v - (v / sel.mper) * sel.mper;
  if (!vec_selected(sel, inst, col)) return;
  const int ix0 = blockIdx.x * TX;
  for (int t = threadIdx.x; t < ny; t += blockDim.x) tw[t] = plan.twy[t];
  cplx<T>* Dv = D + long(v) * 3 * nx * ny;
  for (int e = threadIdx.x; e < ny * TX; e += blockDim.x) {
    const int iy = e / TX;
    const int t = e - iy * TX;
    const long o = (long(iy) * nx + ix0 + t) * 3;
    const cplx<T> d0 = Dv[o], d1 = Dv[o + 1], d2 = Dv[o + 2];
    if (EPS) {
      const T* ep = eps_rowm(plan, inst) + (long(iy) * nx + ix0 + t) * 6;
      const T exx = ep[0], eyy = ep[1], ezz = ep[2], exy = ep[3], exz = ep[4], eyz = ep[5];
      buf[(t * 3 + 0) * ny + iy] = (d0 * exx + d1 * exy + d2 * exz) * scale;
      buf[(t * 3 + 1) * ny + iy] = (d0 * exy + d1 * eyy + d2 * eyz) * scale;
      buf[(t * 3 + 2) * ny + iy] = (d0 * exz + d1 * eyz + d2 * ezz) * scale;
    } else {
      buf[(t * 3 + 0) * ny + iy] = d0 * scale;
      buf[(t * 3 + 1) * ny + iy] = d1 * scale;
      buf[(t * 3 + 2) * ny + iy] = d2 * scale;
    }
  }
  __syncthreads();
  fft_lds<T, +1, VPT>(buf, 3 * TX, ny, 1, plan.fy, tw);
  for (int e = threadIdx.x; e < ny * TX * 3; e += blockDim.x) {
    const int c = e % 3;
    const int q = e / 3;
    const int iy = q / TX;
    const int t = q - iy * TX;
    Dv[(long(iy) * nx + ix0 + t) * 3 + c] = buf[(t * 3 + c) * ny + iy];
  }
}

// Row pass of the inverse path: inverse FFT along x, in place.
template <typename T, int VPT>
__global__ __launch_bounds__(OP_BLOCK) void k_pass3_inv(OpPlan<T> plan, cplx<T>* D, VecSel sel) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
  const int nx = plan.nx, ny = plan.ny, RB = plan.rows_per_block;
  cplx<T>* buf = tw + nx;
  const int v = blockIdx.y;
  const int inst = v / sel.mper, col = v - (v / sel.mper) * sel.mper;
  if (!vec_selected(sel, inst, col)) return;
  const int iy0 = blockIdx.x * RB;
  for (int t = threadIdx.x; t < nx; t += blockDim.x) tw[t] = plan.twx[t];
  cplx<T>* Dv = D + long(v) * 3 * nx * ny;
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
// Register-FFT path: power-of-two nx = GX VX, ny = GY VY; intermediate field [v][c][ix][iy].
// ---------------------------------------------------------------------------------------

// The same basis as basis_at, with one sqrt and one reciprocal square root:
// |n x w| = |w| because n is a unit vector orthogonal to w.
template <typename T>
__device__ __forceinline__ Basis<T> basis_fast(int ix, int iy, int nx, int ny, double invdx, double invdy, T k) {
  const T wx = -T(double(fft_index(ix, nx)) * invdx);
  const T wy = -T(double(fft_index(iy, ny)) * invdy);
  const T rho2 = wx * wx + wy * wy;
  Basis<T> b;
  b.mag = sqrt(rho2 + k * k);
  if (rho2 > T(0)) {
    const T ir = T(1) / sqrt(rho2);
    b.n0 = -wy * ir;
    b.n1 = wx * ir;
  } else {
    b.n0 = T(0);
    b.n1 = T(1);
  }
  b.n2 = T(0);
  const T im = T(1) / b.mag;
  b.m0 = b.n1 * k * im;
  b.m1 = -b.n0 * k * im;
  b.m2 = (b.n0 * wy - b.n1 * wx) * im;
  return b;
}

// pass 1: rows iy0 .. iy0 + RB - 1 of one vector; the LX/8 lanes of group r hold row
// iy0 + r (lane g: ix = g + (LX/8) q, q < 8).  Curl in k-space, FFT along x, then the rows
// are written column-major through an LDS tile (RB consecutive iy per ix: whole lines).
template <typename T, int OP, int LX, int RB>
__global__ __launch_bounds__(RB*(LX / 8)) void k_rpass1(OpPlan<T> plan, BView<T> X, cplx<T>* D, VecSel sel,
                                                         const double* kinst) {
  constexpr int NT = RB * (LX / 8), GX = LX / 8;
  constexpr int SG = reg_fft_scratch<LX>(), TS = RB + 1;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* work = tw + LX;
  const int v = blockIdx.y;
  const int inst = v / sel.mper, col = v - (v / sel.mper) * sel.mper;
  if (!vec_selected(sel, inst, col)) return;
  const int ny = plan.ny;
  for (int t = threadIdx.x; t < LX; t += NT) tw[t] = plan.twx[t];
  const int r = threadIdx.x / GX, g = threadIdx.x - r * GX;
  const int iy0 = blockIdx.x * RB, iy = iy0 + r;
  const T k = T(kinst[inst]);
  const cplx<T>* Xv = X.base + inst * X.istride + col * X.ld;
  cplx<T> d0[8], d1[8], d2[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int ix = g + GX * q;
    const long p = ix + long(LX) * iy;
    const cplx<T> h1 = Xv[2 * p], h2 = Xv[2 * p + 1];
    const Basis<T> b = basis_fast<T>(ix, iy, LX, ny, plan.invdx, plan.invdy, k);
    if (OP == OP_M) {
      d0[q] = (h2 * b.m0 - h1 * b.n0) * b.mag;
      d1[q] = (h2 * b.m1 - h1 * b.n1) * b.mag;
      d2[q] = (h2 * b.m2) * b.mag;
    } else if (OP == OP_P) {
      const T im = T(1) / b.mag;
      d0[q] = (h1 * b.n0 - h2 * b.m0) * im;
      d1[q] = (h1 * b.n1 - h2 * b.m1) * im;
      d2[q] = (h2 * b.m2) * (-im);
    } else {
      d0[q] = T(-1) * (h1 * b.m1 + h2 * b.n1);
      d1[q] = h1 * b.m0 + h2 * b.n0;
      d2[q] = cplx<T>{T(0), T(0)};
    }
  }
  __syncthreads();  // twiddles
  cplx<T>* xs = work + r * SG;
  reg_fft<T, Dirs<OP>::x1, LX>(d0, g, xs, tw);
  reg_fft<T, Dirs<OP>::x1, LX>(d1, g, xs, tw);
  if (OP != OP_MKH) reg_fft<T, Dirs<OP>::x1, LX>(d2, g, xs, tw);
  cplx<T>* Dv = D + long(v) * 3 * LX * ny;
  auto store = [&](const cplx<T>(&dc)[8], int c) {
    __syncthreads();  // FFT scratch and the previous tile are free
#pragma unroll
    for (int q = 0; q < 8; ++q) work[(g + GX * q) * TS + r] = dc[q];
    __syncthreads();
    cplx<T>* Dc = Dv + long(c) * LX * ny + iy0;
#pragma unroll
    for (int u = 0; u < LX * RB / NT; ++u) {
      const int e = threadIdx.x + u * NT;
      const int ix = e / RB, rr = e - (e / RB) * RB;
      Dc[long(ix) * ny + rr] = work[ix * TS + rr];
    }
  };
  store(d0, 0);
  store(d1, 1);
  store(d2, 2);
}

// pass 2: one column ix per group of LY/8 lanes (lane g holds iy = g + (LY/8) q): FFT along
// y, eps^-1 (or eps_ave) per pixel, inverse FFT along y, in place.  No block barrier after
// the twiddles are staged.
template <typename T, int OP, int LY, int NT>
__global__ __launch_bounds__(NT) void k_rpass2(OpPlan<T> plan, cplx<T>* D, VecSel sel) {
  constexpr int GY = LY / 8, CB = NT / GY;
  constexpr int SG = reg_fft_scratch<LY>();
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* work = tw + LY;
  const int v = blockIdx.y;
  const int inst = v / sel.mper, col = v - (v / sel.mper) * sel.mper;
  if (!vec_selected(sel, inst, col)) return;
  const int nx = plan.nx;
  for (int t = threadIdx.x; t < LY; t += NT) tw[t] = plan.twy[t];
  const int cix = threadIdx.x / GY, g = threadIdx.x - cix * GY;
  const int ix = blockIdx.x * CB + cix;
  cplx<T>* D0 = D + long(v) * 3 * nx * LY + long(ix) * LY;
  cplx<T>* D1 = D0 + long(nx) * LY;
  cplx<T>* D2 = D1 + long(nx) * LY;
  cplx<T> a0[8], a1[8], a2[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int iy = g + GY * q;
    a0[q] = D0[iy];
    a1[q] = D1[iy];
    a2[q] = D2[iy];
  }
  __syncthreads();  // twiddles
  cplx<T>* xs = work + cix * SG;
  reg_fft<T, Dirs<OP>::y1, LY>(a0, g, xs, tw);
  reg_fft<T, Dirs<OP>::y1, LY>(a1, g, xs, tw);
  reg_fft<T, Dirs<OP>::y1, LY>(a2, g, xs, tw);
  const T* ep = eps_colm(plan, inst) + long(ix) * LY * 6;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const T* e = ep + (g + GY * q) * 6;
    const cplx<T> x0 = a0[q], x1 = a1[q], x2 = a2[q];
    if (OP == OP_P) {
      const T ea = T(3) / (e[0] + e[1] + e[2]);
      a0[q] = x0 * ea;
      a1[q] = x1 * ea;
      a2[q] = x2 * ea;
    } else {
      const T exx = e[0], eyy = e[1], ezz = e[2], exy = e[3], exz = e[4], eyz = e[5];
      a0[q] = x0 * exx + x1 * exy + x2 * exz;
      a1[q] = x0 * exy + x1 * eyy + x2 * eyz;
      a2[q] = x0 * exz + x1 * eyz + x2 * ezz;
    }
  }
  reg_fft<T, Dirs<OP>::y2, LY>(a0, g, xs, tw);
  reg_fft<T, Dirs<OP>::y2, LY>(a1, g, xs, tw);
  reg_fft<T, Dirs<OP>::y2, LY>(a2, g, xs, tw);
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int iy = g + GY * q;
    D0[iy] = a0[q];
    D1[iy] = a1[q];
    D2[iy] = a2[q];
  }
}

// pass 3: rows iy0 .. iy0 + RB - 1 read column-major through an LDS tile (the next
// component's lines are fetched while the current one is unpacked), inverse FFT along x,
// curl back onto (m, n) scaled by 1/N; for OP_MKH dotted with the input.
template <typename T, int OP, int LX, int RB>
__global__ __launch_bounds__(RB*(LX / 8)) void k_rpass3(OpPlan<T> plan, BView<T> Hin, BView<T> Y, cplx<T>* D,
                                                         VecSel sel, const double* kinst, double* dot_partial,
                                                         int dps) {
  constexpr int NT = RB * (LX / 8), GX = LX / 8;
  constexpr int SG = reg_fft_scratch<LX>(), TS = RB + 1, PER = LX * RB / NT;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* work = tw + LX;
  const int v = blockIdx.y;
  const int inst = v / sel.mper, col = v - (v / sel.mper) * sel.mper;
  if (!vec_selected(sel, inst, col)) return;
  const int ny = plan.ny;
  for (int t = threadIdx.x; t < LX; t += NT) tw[t] = plan.twx[t];
  const int r = threadIdx.x / GX, g = threadIdx.x - r * GX;
  const int iy0 = blockIdx.x * RB, iy = iy0 + r;
  const cplx<T>* Dv = D + long(v) * 3 * LX * ny + iy0;
  cplx<T> e0[8], e1[8], e2[8];
  cplx<T> pre[PER];
  auto fetch = [&](int c) {
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int e = threadIdx.x + u * NT;
      const int ix = e / RB, rr = e - (e / RB) * RB;
      pre[u] = Dv[long(c) * LX * ny + long(ix) * ny + rr];
    }
  };
  auto unpack = [&](cplx<T>(&ec)[8], int cnext) {
    __syncthreads();  // the tile is free
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int e = threadIdx.x + u * NT;
      const int ix = e / RB, rr = e - (e / RB) * RB;
      work[ix * TS + rr] = pre[u];
    }
    if (cnext < 3) fetch(cnext);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 8; ++q) ec[q] = work[(g + GX * q) * TS + r];
  };
  fetch(0);
  unpack(e0, 1);
  unpack(e1, 2);
  unpack(e2, 3);
  __syncthreads();  // tile reads done: the FFT scratch overlays it
  cplx<T>* xs = work + r * SG;
  reg_fft<T, Dirs<OP>::x2, LX>(e0, g, xs, tw);
  reg_fft<T, Dirs<OP>::x2, LX>(e1, g, xs, tw);
  reg_fft<T, Dirs<OP>::x2, LX>(e2, g, xs, tw);
  const T k = T(kinst[inst]);
  const T scale = T(1.0 / (double(LX) * double(ny)));
  double acc = 0.0;
  const cplx<T>* Hv = Hin.base + inst * Hin.istride + col * Hin.ld;
  cplx<T>* Yv = (OP == OP_MKH) ? nullptr : Y.base + inst * Y.istride + col * Y.ld;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int ix = g + GX * q;
    const Basis<T> b = basis_fast<T>(ix, iy, LX, ny, plan.invdx, plan.invdy, k);
    const cplx<T> en = e0[q] * b.n0 + e1[q] * b.n1;
    const cplx<T> em = e0[q] * b.m0 + e1[q] * b.m1 + e2[q] * b.m2;
    cplx<T> h1, h2;
    if (OP == OP_P) {
      const T sc = scale / b.mag;
      h1 = en * sc;
      h2 = em * (-sc);
    } else {
      const T sc = b.mag * scale;
      h1 = en * (-sc);
      h2 = em * sc;
    }
    const long p = ix + long(LX) * iy;
    if (OP == OP_MKH) {
      const cplx<T> a1 = Hv[2 * p], a2 = Hv[2 * p + 1];
      acc += double(a1.x) * double(h1.x) + double(a1.y) * double(h1.y) + double(a2.x) * double(h2.x) +
             double(a2.y) * double(h2.y);
    } else {
      Yv[2 * p] = h1;
      Yv[2 * p + 1] = h2;
    }
  }
  if (OP == OP_MKH) {
    __shared__ double red[NT];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
      if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) dot_partial[long(v) * dps + blockIdx.x] = -red[0];
  }
}

template <typename T, int OP, int LX, int RB>
void rop_rows_rb(int pass, const OpPlan<T>& plan, BView<T> X, BView<T> Y, cplx<T>* D, const VecSel& sel,
                 const double* kinst, double* dot_partial, int dps, hipStream_t stream) {
  constexpr int NT = RB * (LX / 8);
  constexpr int SG = reg_fft_scratch<LX>(), TS = RB + 1;
  const size_t lds = sizeof(cplx<T>) * (LX + std::max(RB * SG, LX * TS));
  dim3 grid(plan.ny / RB, sel.ninst * sel.mper);
  if (pass == 1)
    hipLaunchKernelGGL((k_rpass1<T, OP, LX, RB>), grid, dim3(NT), lds, stream, plan, X, D, sel, kinst);
  else
    hipLaunchKernelGGL((k_rpass3<T, OP, LX, RB>), grid, dim3(NT), lds, stream, plan, X, Y, D, sel, kinst,
                       dot_partial, dps);
}

// Blocks below which a register-path launch takes the halved block shape when the plan
// leaves it open (two blocks per CU of a 256-CU device): launches of few vectors (the
// refinement's correction solves, the shared first eigensolve, the last active bands)
// otherwise leave CUs idle.
constexpr int REG_FILL_BLOCKS = 512;

// rows per block: reg_rows<T>() or half of it (plan.reg_rb; 0: half in pass 1 of M and P
// launches of fewer than REG_FILL_BLOCKS blocks).  Pass 3 keeps the full count when the plan
// leaves it open: its column-major reads of half as many rows are 64-byte runs, which fetch
// whole 128-byte lines (twice the intermediate's bytes, counted); the HMkH launches keep the
// plan's count, which fixes their dot-partial layout.
template <typename T, int OP, int LX>
void rop_rows(int pass, const OpPlan<T>& plan, BView<T> X, BView<T> Y, cplx<T>* D, const VecSel& sel,
              const double* kinst, double* dot_partial, int dps, hipStream_t stream) {
  constexpr int R0 = reg_rows<T>();
  int rb = reg_rows_of(plan);
  if (plan.reg_rb == 0 && pass == 1 && (OP == OP_M || OP == OP_P) &&
      long(plan.ny / R0) * sel.ninst * sel.mper < REG_FILL_BLOCKS)
    rb = R0 / 2;
  if (rb == R0)
    rop_rows_rb<T, OP, LX, R0>(pass, plan, X, Y, D, sel, kinst, dot_partial, dps, stream);
  else if (rb == R0 / 2)
    rop_rows_rb<T, OP, LX, R0 / 2>(pass, plan, X, Y, D, sel, kinst, dot_partial, dps, stream);
  else
    throw Error(OMX_ERR_INTERNAL, "register FFT path: unsupported rows per block");
}

template <typename T, int OP, int LY, int NT>
void rop_cols_nt(const OpPlan<T>& plan, cplx<T>* D, const VecSel& sel, hipStream_t stream) {
  constexpr int CB = NT / (LY / 8);
  const size_t lds = sizeof(cplx<T>) * (LY + CB * reg_fft_scratch<LY>());
  dim3 grid(plan.nx / CB, sel.ninst * sel.mper);
  hipLaunchKernelGGL((k_rpass2<T, OP, LY, NT>), grid, dim3(NT), lds, stream, plan, D, sel);
}

// threads per block of pass 2: REG_COL_THREADS or half of it (plan.reg_nt2; 0: half for
// launches of fewer than REG_FILL_BLOCKS blocks)
template <typename T, int OP, int LY>
void rop_cols(const OpPlan<T>& plan, cplx<T>* D, const VecSel& sel, hipStream_t stream) {
  int nt = plan.reg_nt2 > 0 ? plan.reg_nt2 : REG_COL_THREADS;
  if (plan.reg_nt2 == 0 && long(plan.nx / (REG_COL_THREADS / (LY / 8))) * sel.ninst * sel.mper < REG_FILL_BLOCKS)
    nt = REG_COL_THREADS / 2;
  if (nt == REG_COL_THREADS)
    rop_cols_nt<T, OP, LY, REG_COL_THREADS>(plan, D, sel, stream);
  else if (nt == REG_COL_THREADS / 2)
    rop_cols_nt<T, OP, LY, REG_COL_THREADS / 2>(plan, D, sel, stream);
  else
    throw Error(OMX_ERR_INTERNAL, "register FFT path: unsupported pass-2 block");
}

template <typename T, int OP>
void launch_rop(const OpPlan<T>& plan, BView<T> X, BView<T> Y, cplx<T>* D, const VecSel& sel, const double* kinst,
                double* dot_partial, int dps, hipStream_t stream) {
  auto rows = [&](int pass) {
    switch (plan.nx) {
      case 64: rop_rows<T, OP, 64>(pass, plan, X, Y, D, sel, kinst, dot_partial, dps, stream); break;
      case 128: rop_rows<T, OP, 128>(pass, plan, X, Y, D, sel, kinst, dot_partial, dps, stream); break;
      case 256: rop_rows<T, OP, 256>(pass, plan, X, Y, D, sel, kinst, dot_partial, dps, stream); break;
      case 512: rop_rows<T, OP, 512>(pass, plan, X, Y, D, sel, kinst, dot_partial, dps, stream); break;
      default: throw Error(OMX_ERR_INTERNAL, "register FFT path: unsupported nx");
    }
  };
  rows(1);
  switch (plan.ny) {
    case 64: rop_cols<T, OP, 64>(plan, D, sel, stream); break;
    case 128: rop_cols<T, OP, 128>(plan, D, sel, stream); break;
    case 256: rop_cols<T, OP, 256>(plan, D, sel, stream); break;
    case 512: rop_cols<T, OP, 512>(plan, D, sel, stream); break;
    default: throw Error(OMX_ERR_INTERNAL, "register FFT path: unsupported ny");
  }
  rows(3);
}

template <typename T, int VPT>
void launch_fields_vpt(FieldOp fop, const OpPlan<T>& plan, BView<T> X, cplx<T>* D, const VecSel& sel,
                       const double* kinst, T scale, hipStream_t stream) {
  const int nvec = sel.ninst * sel.mper;
  const size_t cs = sizeof(cplx<T>);
  dim3 g1(plan.ny / plan.rows_per_block, nvec);
  dim3 g2(plan.nx / plan.cols_per_block, nvec);
  const size_t lds13 = cs * (size_t(plan.nx) + size_t(plan.rows_per_block) * 3 * plan.nx);
  const size_t lds2 = cs * (size_t(plan.ny) + size_t(plan.cols_per_block) * 3 * plan.ny);
  if (fop == FIELD_DK || fop == FIELD_DKK || fop == FIELD_H) {
    if (fop == FIELD_DK) {
      hipLaunchKernelGGL((k_pass1<T, OP_MKH, VPT>), g1, dim3(OP_BLOCK), lds13, stream, plan, X, D, sel, kinst);
      scale = -scale;  // d kx_tc / dk = -zx_tc
    } else if (fop == FIELD_H) {
      hipLaunchKernelGGL((k_pass1<T, MAP_TC, VPT>), g1, dim3(OP_BLOCK), lds13, stream, plan, X, D, sel, kinst);
    } else {
      hipLaunchKernelGGL((k_pass1<T, CURL_DKK, VPT>), g1, dim3(OP_BLOCK), lds13, stream, plan, X, D, sel, kinst);
    }
    hipLaunchKernelGGL((k_pass2_fwd<T, VPT, false>), g2, dim3(OP_BLOCK), lds2, stream, plan, D, sel, scale);
  } else if (fop == FIELD_E || fop == FIELD_D) {
    hipLaunchKernelGGL((k_pass1<T, OP_M, VPT>), g1, dim3(OP_BLOCK), lds13, stream, plan, X, D, sel, kinst);
    if (fop == FIELD_E)
      hipLaunchKernelGGL((k_pass2_fwd<T, VPT, true>), g2, dim3(OP_BLOCK), lds2, stream, plan, D, sel, scale);
    else
      hipLaunchKernelGGL((k_pass2_fwd<T, VPT, false>), g2, dim3(OP_BLOCK), lds2, stream, plan, D, sel, scale);
  } else {
    if (fop == FIELD_EINV)
      hipLaunchKernelGGL((k_pass2_inv<T, VPT, true>), g2, dim3(OP_BLOCK), lds2, stream, plan, D, sel, scale);
    else
      hipLaunchKernelGGL((k_pass2_inv<T, VPT, false>), g2, dim3(OP_BLOCK), lds2, stream, plan, D, sel, scale);
    hipLaunchKernelGGL((k_pass3_inv<T, VPT>), g1, dim3(OP_BLOCK), lds13, stream, plan, D, sel);
  }
}

template <typename T, int OP, int VPT>
void launch_op_vpt(const OpPlan<T>& plan, BView<T> X, BView<T> Y, cplx<T>* D, const VecSel& sel,
                   const double* kinst, double* dot_partial, int dps, hipStream_t stream) {
  const int nvec = sel.ninst * sel.mper;
  const size_t cs = sizeof(cplx<T>);
  dim3 g1(plan.ny / plan.rows_per_block, nvec);
  dim3 g2(plan.nx / plan.cols_per_block, nvec);
  const size_t lds13 = cs * (size_t(plan.nx) + size_t(plan.rows_per_block) * 3 * plan.nx);
  const size_t lds2 = cs * (size_t(plan.ny) + size_t(plan.cols_per_block) * 3 * plan.ny);
  hipLaunchKernelGGL((k_pass1<T, OP, VPT>), g1, dim3(OP_BLOCK), lds13, stream, plan, X, D, sel, kinst);
  hipLaunchKernelGGL((k_pass2<T, OP, VPT>), g2, dim3(OP_BLOCK), lds2, stream, plan, D, sel);
  hipLaunchKernelGGL((k_pass3<T, OP, VPT>), g1, dim3(OP_BLOCK), lds13, stream, plan, X, Y, D, sel, kinst,
                     dot_partial, dps);
}

template <typename T, int OP>
void launch_op(const OpPlan<T>& plan, BView<T> X, BView<T> Y, cplx<T>* D, const VecSel& sel,
               const double* kinst, double* dot_partial, int dps, hipStream_t stream) {
  if (plan.reg) {
    launch_rop<T, OP>(plan, X, Y, D, sel, kinst, dot_partial, dps, stream);
    return;
  }
  switch (plan.vpt) {
    case 8: launch_op_vpt<T, OP, 8>(plan, X, Y, D, sel, kinst, dot_partial, dps, stream); break;
    case 16: launch_op_vpt<T, OP, 16>(plan, X, Y, D, sel, kinst, dot_partial, dps, stream); break;
    case 32: launch_op_vpt<T, OP, 32>(plan, X, Y, D, sel, kinst, dot_partial, dps, stream); break;
    default: throw Error(OMX_ERR_INTERNAL, "operator plan has an unsupported values-per-thread bound");
  }
}

}  // namespace

template <typename T>
void launch_operator(OpKind op, const OpPlan<T>& plan, BView<T> X, BView<T> Y, cplx<T>* D, const VecSel& sel,
                     const double* kinst, double* dot_partial, int dps, hipStream_t stream) {
  if (plan.nz > 1) {
    launch_operator3d<T>(op, plan, X, Y, D, sel, kinst, dot_partial, dps, stream);
    return;
  }
  switch (op) {
    case OP_M: launch_op<T, OP_M>(plan, X, Y, D, sel, kinst, dot_partial, dps, stream); break;
    case OP_P: launch_op<T, OP_P>(plan, X, Y, D, sel, kinst, dot_partial, dps, stream); break;
    case OP_MKH: launch_op<T, OP_MKH>(plan, X, Y, D, sel, kinst, dot_partial, dps, stream); break;
    default: throw Error(OMX_ERR_UNSUPPORTED, "the period-weighted MkH needs a 3D cell");
  }
  OMX_HIP_CHECK(hipGetLastError());
}

template <typename T>
void launch_fields(FieldOp fop, const OpPlan<T>& plan, BView<T> X, cplx<T>* D, const VecSel& sel,
                   const double* kinst, double scale, hipStream_t stream) {
  if (plan.nz > 1) {
    launch_fields3d<T>(fop, plan, X, D, sel, kinst, scale, stream);
    return;
  }
  switch (plan.vpt) {
    case 8: launch_fields_vpt<T, 8>(fop, plan, X, D, sel, kinst, T(scale), stream); break;
    case 16: launch_fields_vpt<T, 16>(fop, plan, X, D, sel, kinst, T(scale), stream); break;
    case 32: launch_fields_vpt<T, 32>(fop, plan, X, D, sel, kinst, T(scale), stream); break;
    default: throw Error(OMX_ERR_INTERNAL, "operator plan has an unsupported values-per-thread bound");
  }
  OMX_HIP_CHECK(hipGetLastError());
}

template void launch_fields<double>(FieldOp, const OpPlan<double>&, BView<double>, cplx<double>*, const VecSel&,
                                    const double*, double, hipStream_t);
template void launch_operator<float>(OpKind, const OpPlan<float>&, BView<float>, BView<float>, cplx<float>*,
                                     const VecSel&, const double*, double*, int, hipStream_t);
template void launch_operator<double>(OpKind, const OpPlan<double>&, BView<double>, BView<double>,
                                      cplx<double>*, const VecSel&, const double*, double*, int, hipStream_t);

bool fft_factor_radix4(int L, FFTFactors& f) {
  int rem = L, p2 = 0;
  while (rem % 2 == 0) {
    rem /= 2;
    ++p2;
  }
  if (p2 < 2 || p2 % 2 != 0 || p2 / 2 > FFT_MAX_RADICES) return fft_factor(L, f);
  int others[FFT_MAX_RADICES];
  int no = 0;
  for (int r : {3, 5}) {
    while (rem % r == 0) {
      if (no >= FFT_MAX_RADICES) return fft_factor(L, f);
      others[no++] = r;
      rem /= r;
    }
  }
  if (rem != 1 || p2 / 2 + no > FFT_MAX_RADICES) return fft_factor(L, f);
  f.L = L;
  f.n = 0;
  for (int i = 0; i < p2 / 2; ++i) f.r[f.n++] = 4;
  for (int i = 0; i < no; ++i) f.r[f.n++] = others[i];
  return true;
}

bool fft_factor(int L, FFTFactors& f) {
  f.L = L;
  f.n = 0;
  int rem = L;
  int p2 = 0;
  while (rem % 2 == 0) {
    rem /= 2;
    ++p2;
  }
  int others[FFT_MAX_RADICES];
  int no = 0;
  for (int r : {3, 5}) {
    while (rem % r == 0) {
      if (no >= FFT_MAX_RADICES) return false;
      others[no++] = r;
      rem /= r;
    }
  }
  if (rem != 1) return false;
  int pow2[FFT_MAX_RADICES];
  int np = 0;
  int e = p2;
  if (e == 1) {
    pow2[np++] = 2;
    e = 0;
  } else if (e % 3 == 1) {
    pow2[np++] = 4;
    pow2[np++] = 4;
    e -= 4;
  } else if (e % 3 == 2) {
    pow2[np++] = 4;
    e -= 2;
  }
  while (e > 0) {
    pow2[np++] = 8;
    e -= 3;
  }
  if (np + no > FFT_MAX_RADICES) return false;
  // larger radices first
  for (int i = np - 1; i >= 0; --i) f.r[f.n++] = pow2[i];
  for (int i = 0; i < no; ++i) f.r[f.n++] = others[i];
  return true;
}

}  // namespace omx
