// Register-resident FFTs of power-of-two length L (64 .. 512) for the row / column passes.
//
// A sequence is held by a group of G = L / 8 consecutive lanes of one wavefront: lane g
// holds x[g + G q] in v[q], q < 8.  The transform is a Stockham FFT whose passes run in
// registers: pass P (radix R, NS = product of the earlier radices) has butterflies
// j = g + G b, b < 8 / R, each taking x[j + r L/R], r < R, twiddled by W_{NS R}^{(j mod NS) r},
// and writing its outputs to (j / NS) NS R + (j mod NS) + r NS.  Between passes the group
// exchanges values through LDS (padded index i + i/8 against bank conflicts) under
// wavefront-scope ordering only; the first pass reads registers directly and the last
// pass's outputs are already in the input layout (lane g holds X[g + G q]).  Radices:
// 64 = 8 8, 128 = 8 8 2, 256 = 8 8 4, 512 = 8 8 8 (one or two LDS exchanges).
//
// The 8-, 4- and 2-point kernels come from fft_lds.hpp; the 16- and 32-point in-register
// DFTs below (built by the same splitting) serve other callers.
#pragma once
#include "omx_common.hpp"
#include "fft_lds.hpp"

namespace omx {

// cos(2 pi r / 32), r = 0..8 (first octant; the rest by symmetry), rounded from extended
// precision
__host__ __device__ constexpr double tw32q(int r) {
  return r == 0 ? 1.0
       : r == 1 ? 0.9807852804032304
       : r == 2 ? 0.9238795325112867
       : r == 3 ? 0.8314696123025452
       : r == 4 ? 0.7071067811865476
       : r == 5 ? 0.5555702330196022
       : r == 6 ? 0.3826834323650898
       : r == 7 ? 0.19509032201612828
       : 0.0;
}
// cos and sin of 2 pi m / 32 for any integer m
__host__ __device__ constexpr double tw32c(int m) {
  return ((m % 32 + 32) % 32) / 8 == 0 ? tw32q((m % 32 + 32) % 32 % 8)
       : ((m % 32 + 32) % 32) / 8 == 1 ? -tw32q(8 - (m % 32 + 32) % 32 % 8)
       : ((m % 32 + 32) % 32) / 8 == 2 ? -tw32q((m % 32 + 32) % 32 % 8)
       : tw32q(8 - (m % 32 + 32) % 32 % 8);
}
__host__ __device__ constexpr double tw32s(int m) { return tw32c(m - 8); }

// a * exp(DIR 2 pi i m / N) for N dividing 32
template <int N, int DIR, typename T> __device__ __forceinline__ cplx<T> rot(cplx<T> a, int m) {
  const int mm = m * (32 / N);
  const T c = T(tw32c(mm));
  const T s = DIR < 0 ? T(-tw32s(mm)) : T(tw32s(mm));
  return cplx<T>{a.x * c - a.y * s, a.x * s + a.y * c};
}

template <int N, int DIR, typename T> struct RDft;
template <int DIR, typename T> struct RDft<2, DIR, T> {
  static __device__ __forceinline__ void run(cplx<T>* v) { Dft<2, DIR, T>::run(v); }
};
template <int DIR, typename T> struct RDft<4, DIR, T> {
  static __device__ __forceinline__ void run(cplx<T>* v) { Dft<4, DIR, T>::run(v); }
};
template <int DIR, typename T> struct RDft<8, DIR, T> {
  static __device__ __forceinline__ void run(cplx<T>* v) { Dft<8, DIR, T>::run(v); }
};
// N = N1 N2 with an N1-point inner DFT over n1 (stride N2) and an N2-point outer DFT
template <int N1, int N2, int DIR, typename T> struct RDftSplit {
  static __device__ __forceinline__ void run(cplx<T>* v) {
    constexpr int N = N1 * N2;
    cplx<T> a[N2][N1];
#pragma unroll
    for (int n2 = 0; n2 < N2; ++n2) {
#pragma unroll
      for (int n1 = 0; n1 < N1; ++n1) a[n2][n1] = v[N2 * n1 + n2];
      RDft<N1, DIR, T>::run(a[n2]);
#pragma unroll
      for (int k1 = 1; k1 < N1; ++k1)
        if (n2 > 0) a[n2][k1] = rot<N, DIR, T>(a[n2][k1], n2 * k1);
    }
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) {
      cplx<T> b[N2];
#pragma unroll
      for (int n2 = 0; n2 < N2; ++n2) b[n2] = a[n2][k1];
      RDft<N2, DIR, T>::run(b);
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2) v[k1 + N1 * k2] = b[k2];
    }
  }
};
template <int DIR, typename T> struct RDft<16, DIR, T> {
  static __device__ __forceinline__ void run(cplx<T>* v) { RDftSplit<4, 4, DIR, T>::run(v); }
};
template <int DIR, typename T> struct RDft<32, DIR, T> {
  static __device__ __forceinline__ void run(cplx<T>* v) { RDftSplit<8, 4, DIR, T>::run(v); }
};

// order LDS accesses of one wavefront (the group never spans wavefronts)
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// LDS index of sequence element i in a group's scratch
__device__ __forceinline__ int reg_pad(int i) { return i + (i >> 3); }
// complex values of LDS scratch one group needs
template <int L> constexpr int reg_fft_scratch() { return L + L / 8 + 8; }

__host__ __device__ constexpr int reg_npass(int L) { return L == 64 ? 2 : (L == 128 || L == 256 || L == 512) ? 3 : 0; }
__host__ __device__ constexpr int reg_radix(int L, int P) { return P < 2 ? 8 : (L == 128 ? 2 : L == 256 ? 4 : 8); }
__host__ __device__ constexpr int reg_ns(int P) { return P == 0 ? 1 : P == 1 ? 8 : 64; }
// lengths the register FFT supports
__host__ __device__ constexpr bool reg_fft_ok(int L) { return L == 64 || L == 128 || L == 256 || L == 512; }

template <typename T, int DIR, int L, int P>
__device__ __forceinline__ void reg_fft_pass(cplx<T> (&v)[8], int g, cplx<T>* xs, const cplx<T>* twL) {
  constexpr int NP = reg_npass(L), R = reg_radix(L, P), NS = reg_ns(P), G = L / 8, NB = 8 / R;
  if constexpr (P > 0) {
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int r = 0; r < R; ++r) v[b * R + r] = xs[reg_pad(g + G * b + r * (L / R))];
  }
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    if constexpr (NS > 1) {
      const int k = (g + G * b) % NS;
#pragma unroll
      for (int r = 1; r < R; ++r) {
        cplx<T> w = twL[k * r * (L / (NS * R))];
        if (DIR > 0) w.y = -w.y;
        v[b * R + r] = v[b * R + r] * w;
      }
    }
    RDft<R, DIR, T>::run(&v[b * R]);
  }
  if constexpr (P < NP - 1) {
    wave_lds_sync();  // earlier reads of the scratch are complete
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int j = g + G * b, k = j % NS;
#pragma unroll
      for (int r = 0; r < R; ++r) xs[reg_pad((j / NS) * NS * R + k + r * NS)] = v[b * R + r];
    }
    wave_lds_sync();
    reg_fft_pass<T, DIR, L, P + 1>(v, g, xs, twL);
  } else if constexpr (NB > 1) {
    cplx<T> t[8];
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int r = 0; r < R; ++r) t[b + r * NB] = v[b * R + r];
#pragma unroll
    for (int q = 0; q < 8; ++q) v[q] = t[q];
  }
}

// Transform the sequence held by this lane's group (g: lane index within the group);
// xs: the group's LDS scratch (reg_fft_scratch<L>() values); twL[t] = exp(-2 pi i t / L).
template <typename T, int DIR, int L>
__device__ __forceinline__ void reg_fft(cplx<T> (&v)[8], int g, cplx<T>* xs, const cplx<T>* twL) {
  static_assert(reg_fft_ok(L), "unsupported register FFT length");
  reg_fft_pass<T, DIR, L, 0>(v, g, xs, twL);
}

}  // namespace omx
