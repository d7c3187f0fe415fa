// Kernels of the FP64 eigenpair refinement (see refine.hpp).  All are HBM-bound streaming
// passes over a few columns of K complex FP64 values per instance.
#include "refine.hpp"

#include <algorithm>

namespace omx {

namespace {

constexpr int RB = 256;
constexpr int RQ = 32;  // rows staged per chunk in the Gram kernel

__device__ __forceinline__ const c64* col_ptr(const RCols& C, int i, int c, int n1) {
  return c < n1 ? C.blk + i * C.istride + c * C.ld : C.extra + i * C.ld;
}

// Partials of U^H V (and, when part_uu is given, of U^H U from the same staged rows: the
// Rayleigh-Ritz pair of Gram matrices in one pass over the block)
__global__ __launch_bounds__(RB) void k_bgram(RCols U, RCols V, int vmode, const int* n1, long K, c64* part,
                                              c64* part_uu) {
  __shared__ c64 us[RF_MAXC][RQ + 1], vs[RF_MAXC][RQ + 1];
  const int i = blockIdx.y;
  const int m = n1[i];
  const int nu = m + (U.extra ? 1 : 0);
  const int nv1 = vmode == 0 ? m : 0;
  const int nv = nv1 + (V.extra ? 1 : 0);
  const int t = threadIdx.x;
  double ar[2] = {0, 0}, ai[2] = {0, 0}, br[2] = {0, 0}, bi[2] = {0, 0};
  if (m > 0) {
    // a chunk's entries are fetched into registers one chunk ahead (every load of a chunk in
    // flight together, and during the previous chunk's products), then staged in LDS.  Entries
    // outside the columns or rows load a valid address (row 0 of the first column) and are
    // replaced by zero.
    constexpr int NE = (RF_MAXC * RQ + RB - 1) / RB;
    const c64* u0 = U.blk + i * U.istride;
    const c64* v0 = nv1 > 0 ? V.blk + i * V.istride : V.extra + i * V.ld;
    c64 ur[NE], vr[NE];
    bool ok_u[NE], ok_v[NE];
    const long step = long(gridDim.x) * RQ;
    long r0 = long(blockIdx.x) * RQ;
    auto fetch = [&](long rb) {
#pragma unroll
      for (int k = 0; k < NE; ++k) {
        const int e = t + k * RB, c = e / RQ, q = e % RQ;
        const long r = rb + q;
        ok_u[k] = e < RF_MAXC * RQ && c < nu && r < K;
        ok_v[k] = e < RF_MAXC * RQ && c < nv && r < K;
        ur[k] = *(ok_u[k] ? col_ptr(U, i, c, m) + r : u0);
        vr[k] = *(ok_v[k] ? col_ptr(V, i, c, nv1) + r : v0);
      }
    };
    fetch(r0);
    for (; r0 < K; r0 += step) {
#pragma unroll
      for (int k = 0; k < NE; ++k) {
        const int e = t + k * RB;
        if (e < RF_MAXC * RQ) {
          us[e / RQ][e % RQ] = ok_u[k] ? ur[k] : c64{0, 0};
          vs[e / RQ][e % RQ] = ok_v[k] ? vr[k] : c64{0, 0};
        }
      }
      __syncthreads();
      if (r0 + step < K) fetch(r0 + step);
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        const int p = t + s * RB;
        if (p < nu * nv) {
          const int a = p % nu, b = p / nu;
          double sr = 0, si = 0;
#pragma unroll 8
          for (int q = 0; q < RQ; ++q) {
            const c64 x = us[a][q], y = vs[b][q];
            sr += x.x * y.x + x.y * y.y;
            si += x.x * y.y - x.y * y.x;
          }
          ar[s] += sr;
          ai[s] += si;
        }
        if (part_uu && p < nu * nu) {
          const int a = p % nu, b = p / nu;
          double sr = 0, si = 0;
#pragma unroll 8
          for (int q = 0; q < RQ; ++q) {
            const c64 x = us[a][q], y = us[b][q];
            sr += x.x * y.x + x.y * y.y;
            si += x.x * y.y - x.y * y.x;
          }
          br[s] += sr;
          bi[s] += si;
        }
      }
      __syncthreads();
    }
  }
  c64* pp = part + (long(i) * gridDim.x + blockIdx.x) * RF_PAIRS;
  c64* pu = part_uu ? part_uu + (long(i) * gridDim.x + blockIdx.x) * RF_PAIRS : nullptr;
  for (int s = 0; s < 2; ++s) {
    const int p = t + s * RB;
    // store in G's layout a + RF_MAXC b
    if (p < nu * nv) {
      const int a = p % nu, b = p / nu;
      pp[a + RF_MAXC * b] = c64{ar[s], ai[s]};
    }
    if (pu && p < nu * nu) {
      const int a = p % nu, b = p / nu;
      pu[a + RF_MAXC * b] = c64{br[s], bi[s]};
    }
  }
}

// G entry p of instance i = the sum of its nblk partials: one wavefront per entry, lane l
// summing partials l, l + 64, ... and a fixed shuffle tree (the same order on every run)
__global__ __launch_bounds__(RB) void k_bgram_reduce(const c64* part, int nblk, const int* n1, int has_u2, int vmode,
                                                     int has_v2, c64* G) {
  const int i = blockIdx.x;
  const int m = n1[i];
  const int nu = m + has_u2, nv = (vmode == 0 ? m : 0) + has_v2;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int p = w; p < RF_PAIRS; p += RB / 64) {
    const int a = p % RF_MAXC, b = p / RF_MAXC;
    double sr = 0, si = 0;
    if (m > 0 && a < nu && b < nv)
      for (int k = lane; k < nblk; k += 64) {
        const c64 v = part[(long(i) * nblk + k) * RF_PAIRS + p];
        sr += v.x;
        si += v.y;
      }
    for (int off = 32; off > 0; off >>= 1) {
      sr += __shfl_down(sr, off, 64);
      si += __shfl_down(si, off, 64);
    }
    if (lane == 0) G[long(i) * RF_PAIRS + p] = c64{sr, si};
  }
}

__global__ __launch_bounds__(RB) void k_bcombine(RCols U, const c64* C, const int* n1, int omode, c64* out,
                                                 long ostride, long K) {
  const int i = blockIdx.y;
  const int m = n1[i];
  if (m == 0) return;
  const int nin = m + (U.extra ? 1 : 0);
  const int nout = omode == 0 ? m : 1;
  const c64* Ci = C + long(i) * RF_PAIRS;
  for (long r = long(blockIdx.x) * RB + threadIdx.x; r < K; r += long(gridDim.x) * RB) {
    c64 in[RF_MAXC];
#pragma unroll
    for (int l = 0; l < RF_MAXC; ++l) in[l] = l < nin ? col_ptr(U, i, l, m)[r] : c64{0, 0};
    for (int j = 0; j < nout; ++j) {
      double sr = 0, si = 0;
#pragma unroll
      for (int l = 0; l < RF_MAXC; ++l) {
        if (l < nin) {
          const c64 c = Ci[l + RF_MAXC * j];
          sr += in[l].x * c.x - in[l].y * c.y;
          si += in[l].x * c.y + in[l].y * c.x;
        }
      }
      c64* o = omode == 0 ? out + i * ostride + j * U.ld : out + i * U.ld;
      o[r] = c64{sr, si};
    }
  }
}

__global__ __launch_bounds__(RB) void k_band_resid(const c64* X, const c64* AX, long ld, long istride, const int* col,
                                                   const double* lam, const int* n1, long K, c64* xb, c64* rb,
                                                   double* part) {
  const int i = blockIdx.y;
  __shared__ double red[RB];
  double acc = 0;
  if (n1[i] > 0) {
    const c64* x = X + i * istride + long(col[i]) * ld;
    const c64* ax = AX + i * istride + long(col[i]) * ld;
    const double l = lam[i];
    for (long r = long(blockIdx.x) * RB + threadIdx.x; r < K; r += long(gridDim.x) * RB) {
      const c64 a = x[r], b = ax[r];
      const c64 res{b.x - l * a.x, b.y - l * a.y};
      xb[i * ld + r] = a;
      rb[i * ld + r] = res;
      acc += res.x * res.x + res.y * res.y;
    }
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = RB / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[long(i) * gridDim.x + blockIdx.x] = red[0];
}

__global__ __launch_bounds__(RB) void k_scale_cols(c64* x, long ld, const double* scale, const int* n1, long K) {
  const int i = blockIdx.y;
  if (n1[i] == 0) return;
  const double f = scale[i];
  for (long r = long(blockIdx.x) * RB + threadIdx.x; r < K; r += long(gridDim.x) * RB) {
    c64& v = x[i * ld + r];
    v = c64{v.x * f, v.y * f};
  }
}

int rows_grid(long K) { return int(std::min<long>(256, (K + RB - 1) / RB)); }

}  // namespace

void launch_bgram(const RCols& U, const RCols& V, int vmode, const int* n1, int ninst, long K, int nblk, c64* part,
                  c64* G, hipStream_t s, c64* part_uu, c64* B) {
  hipLaunchKernelGGL(k_bgram, dim3(nblk, ninst), dim3(RB), 0, s, U, V, vmode, n1, K, part, B ? part_uu : nullptr);
  hipLaunchKernelGGL(k_bgram_reduce, dim3(ninst), dim3(RB), 0, s, part, nblk, n1, U.extra ? 1 : 0, vmode,
                     V.extra ? 1 : 0, G);
  if (B)
    hipLaunchKernelGGL(k_bgram_reduce, dim3(ninst), dim3(RB), 0, s, part_uu, nblk, n1, U.extra ? 1 : 0, 0,
                       U.extra ? 1 : 0, B);
  OMX_HIP_CHECK(hipGetLastError());
}

void launch_bcombine(const RCols& U, const c64* C, const int* n1, int omode, c64* out, long ostride, int ninst,
                     long K, hipStream_t s) {
  hipLaunchKernelGGL(k_bcombine, dim3(rows_grid(K), ninst), dim3(RB), 0, s, U, C, n1, omode, out, ostride, K);
  OMX_HIP_CHECK(hipGetLastError());
}

void launch_band_resid(const c64* X, const c64* AX, long ld, long istride, const int* col, const double* lam,
                       const int* n1, int ninst, long K, int nblk, c64* xb, c64* rb, double* part, hipStream_t s) {
  hipLaunchKernelGGL(k_band_resid, dim3(nblk, ninst), dim3(RB), 0, s, X, AX, ld, istride, col, lam, n1, K, xb, rb,
                     part);
  OMX_HIP_CHECK(hipGetLastError());
}

void launch_scale_cols(c64* x, long ld, const double* scale, const int* n1, int ninst, long K, hipStream_t s) {
  hipLaunchKernelGGL(k_scale_cols, dim3(rows_grid(K), ninst), dim3(RB), 0, s, x, ld, scale, n1, K);
  OMX_HIP_CHECK(hipGetLastError());
}

}  // namespace omx
