// Batched mixed-radix (2,3,4,5,8) Stockham FFT of short 1D transforms held in LDS.
//
// One workgroup transforms `nb` sequences of length L, element i of sequence b at
// buf[b*sb + i*se].  Each pass with radix R runs nb*L/R butterflies; a thread loads the
// R inputs of each of its butterflies, twiddles them, applies the radix-R DFT in
// registers, waits for the whole workgroup, then stores the outputs (Stockham
// auto-sort: no bit reversal).  DIR = -1 is the forward transform (exp(-2 pi i ...),
// FFTW `fft`), DIR = +1 the unnormalised backward transform (FFTW `bfft`).
//
// Twiddles come from a table tw[t] = exp(-2 pi i t / L), t < L, computed on the host in
// extended precision and staged in LDS by the caller.
#pragma once
#include "omx_common.hpp"

namespace omx {

constexpr int FFT_MAX_RADICES = 8;

struct FFTFactors {
  int L;
  int n;
  int r[FFT_MAX_RADICES];
};

template <typename T> struct RadixConst;
template <> struct RadixConst<float> {
  static constexpr float s2 = 0.70710678118654752f;   // 1/sqrt(2)
  static constexpr float s3 = 0.86602540378443865f;   // sqrt(3)/2
  static constexpr float c51 = 0.30901699437494742f;  // cos(2 pi/5)
  static constexpr float c52 = -0.80901699437494742f; // cos(4 pi/5)
  static constexpr float s51 = 0.95105651629515357f;  // sin(2 pi/5)
  static constexpr float s52 = 0.58778525229247313f;  // sin(4 pi/5)
};
template <> struct RadixConst<double> {
  static constexpr double s2 = 0.70710678118654752440;
  static constexpr double s3 = 0.86602540378443864676;
  static constexpr double c51 = 0.30901699437494742410;
  static constexpr double c52 = -0.80901699437494742410;
  static constexpr double s51 = 0.95105651629515357212;
  static constexpr double s52 = 0.58778525229247312917;
};

// multiply by i*DIR
template <int DIR, typename T> __device__ __forceinline__ cplx<T> mul_is(cplx<T> a) {
  return DIR < 0 ? cplx<T>{a.y, -a.x} : cplx<T>{-a.y, a.x};
}

template <int R, int DIR, typename T> struct Dft;

template <int DIR, typename T> struct Dft<2, DIR, T> {
  static __device__ __forceinline__ void run(cplx<T>* v) {
    cplx<T> a = v[0], b = v[1];
    v[0] = a + b;
    v[1] = a - b;
  }
};

template <int DIR, typename T> struct Dft<4, DIR, T> {
  static __device__ __forceinline__ void run(cplx<T>* v) {
    cplx<T> t0 = v[0] + v[2], t1 = v[0] - v[2];
    cplx<T> t2 = v[1] + v[3], t3 = mul_is<DIR>(v[1] - v[3]);
    v[0] = t0 + t2;
    v[2] = t0 - t2;
    v[1] = t1 + t3;
    v[3] = t1 - t3;
  }
};

template <int DIR, typename T> struct Dft<8, DIR, T> {
  static __device__ __forceinline__ void run(cplx<T>* v) {
    cplx<T> e[4] = {v[0], v[2], v[4], v[6]};
    cplx<T> o[4] = {v[1], v[3], v[5], v[7]};
    Dft<4, DIR, T>::run(e);
    Dft<4, DIR, T>::run(o);
    const T h = RadixConst<T>::s2;
    // w8^q = exp(DIR * 2 pi i q / 8)
    cplx<T> o1 = o[1], o3 = o[3];
    cplx<T> w1o = DIR < 0 ? cplx<T>{h * (o1.x + o1.y), h * (o1.y - o1.x)} : cplx<T>{h * (o1.x - o1.y), h * (o1.x + o1.y)};
    cplx<T> w2o = mul_is<DIR>(o[2]);
    cplx<T> w3o = DIR < 0 ? cplx<T>{h * (-o3.x + o3.y), h * (-o3.y - o3.x)} : cplx<T>{h * (-o3.x - o3.y), h * (o3.x - o3.y)};
    v[0] = e[0] + o[0];
    v[4] = e[0] - o[0];
    v[1] = e[1] + w1o;
    v[5] = e[1] - w1o;
    v[2] = e[2] + w2o;
    v[6] = e[2] - w2o;
    v[3] = e[3] + w3o;
    v[7] = e[3] - w3o;
  }
};

template <int DIR, typename T> struct Dft<3, DIR, T> {
  static __device__ __forceinline__ void run(cplx<T>* v) {
    cplx<T> t = v[1] + v[2];
    cplx<T> u = mul_is<DIR>(v[1] - v[2]) * RadixConst<T>::s3;
    cplx<T> a = v[0] - t * T(0.5);
    v[0] = v[0] + t;
    v[1] = a + u;
    v[2] = a - u;
  }
};

template <int DIR, typename T> struct Dft<5, DIR, T> {
  static __device__ __forceinline__ void run(cplx<T>* v) {
    using C = RadixConst<T>;
    cplx<T> t1 = v[1] + v[4], t2 = v[2] + v[3];
    cplx<T> u1 = v[1] - v[4], u2 = v[2] - v[3];
    cplx<T> a1 = v[0] + t1 * C::c51 + t2 * C::c52;
    cplx<T> a2 = v[0] + t1 * C::c52 + t2 * C::c51;
    cplx<T> b1 = mul_is<DIR>(u1 * C::s51 + u2 * C::s52);
    cplx<T> b2 = mul_is<DIR>(u1 * C::s52 - u2 * C::s51);
    v[0] = v[0] + t1 + t2;
    v[1] = a1 + b1;
    v[4] = a1 - b1;
    v[2] = a2 + b2;
    v[3] = a2 - b2;
  }
};

// One Stockham pass of radix R over all nb sequences.  VPT bounds the complex values a
// thread holds between its loads and its stores; the host guarantees
// ceil(nb*L/R / blockDim) <= ceil(VPT/R).
template <typename T, int DIR, int R, int VPT>
__device__ __forceinline__ void fft_pass(cplx<T>* buf, int nb, int L, int sb, int se, int Ns,
                                         const cplx<T>* tw) {
  constexpr int MB = (VPT + R - 1) / R;
  const int T_ = L / R;
  const int nbf = nb * T_;
  const int tstride = L / (Ns * R);
  cplx<T> v[MB][R];
  int outidx[MB];
#pragma unroll
  for (int u = 0; u < MB; ++u) {
    const int bf = threadIdx.x + u * blockDim.x;
    outidx[u] = -1;
    if (bf < nbf) {
      const int b = bf / T_;
      const int j = bf - b * T_;
      const int jq = j / Ns;
      const int k = j - jq * Ns;
      const cplx<T>* base = buf + b * sb;
#pragma unroll
      for (int r = 0; r < R; ++r) v[u][r] = base[(j + r * T_) * se];
      const int ts = k * tstride;
#pragma unroll
      for (int r = 1; r < R; ++r) {
        cplx<T> w = tw[ts * r];
        if (DIR > 0) w.y = -w.y;
        v[u][r] = v[u][r] * w;
      }
      Dft<R, DIR, T>::run(v[u]);
      outidx[u] = b * sb + (jq * Ns * R + k) * se;
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < MB; ++u) {
    if (outidx[u] >= 0) {
#pragma unroll
      for (int q = 0; q < R; ++q) buf[outidx[u] + q * Ns * se] = v[u][q];
    }
  }
  __syncthreads();
}

// Full transform: all passes of the factorisation.  Caller must __syncthreads() after
// filling `buf`; on return `buf` holds the result and all threads are synchronised.
template <typename T, int DIR, int VPT>
__device__ void fft_lds(cplx<T>* buf, int nb, int sb, int se, const FFTFactors& f, const cplx<T>* tw) {
  int Ns = 1;
  for (int p = 0; p < f.n; ++p) {
    const int R = f.r[p];
    switch (R) {
      case 8: fft_pass<T, DIR, 8, VPT>(buf, nb, f.L, sb, se, Ns, tw); break;
      case 4: fft_pass<T, DIR, 4, VPT>(buf, nb, f.L, sb, se, Ns, tw); break;
      case 2: fft_pass<T, DIR, 2, VPT>(buf, nb, f.L, sb, se, Ns, tw); break;
      case 3: fft_pass<T, DIR, 3, VPT>(buf, nb, f.L, sb, se, Ns, tw); break;
      case 5: fft_pass<T, DIR, 5, VPT>(buf, nb, f.L, sb, se, Ns, tw); break;
      default: break;
    }
    Ns *= R;
  }
}

// The same transform over nb sequences with a two-level start: sequence b starts at
// buf + (b / nb2) sb1 + (b % nb2) sb2 (element stride se).  One call then covers, e.g.,
// every z-pencil of a [group][iz][iy] slab (nb2 = ny, sb1 = ny nz, sb2 = 1, se = ny).
template <typename T, int DIR, int R, int VPT>
__device__ __forceinline__ void fft_pass2(cplx<T>* buf, int nb, int nb2, int sb1, int sb2, int L, int se, int Ns,
                                          const cplx<T>* tw) {
  constexpr int MB = (VPT + R - 1) / R;
  const int T_ = L / R;
  const int nbf = nb * T_;
  const int tstride = L / (Ns * R);
  cplx<T> v[MB][R];
  int outidx[MB];
#pragma unroll
  for (int u = 0; u < MB; ++u) {
    const int bf = threadIdx.x + u * blockDim.x;
    outidx[u] = -1;
    if (bf < nbf) {
      const int b = bf / T_;
      const int j = bf - b * T_;
      const int jq = j / Ns;
      const int k = j - jq * Ns;
      const int b1 = b / nb2;
      const int start = b1 * sb1 + (b - b1 * nb2) * sb2;
      const cplx<T>* base = buf + start;
#pragma unroll
      for (int r = 0; r < R; ++r) v[u][r] = base[(j + r * T_) * se];
      const int ts = k * tstride;
#pragma unroll
      for (int r = 1; r < R; ++r) {
        cplx<T> w = tw[ts * r];
        if (DIR > 0) w.y = -w.y;
        v[u][r] = v[u][r] * w;
      }
      Dft<R, DIR, T>::run(v[u]);
      outidx[u] = start + (jq * Ns * R + k) * se;
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < MB; ++u) {
    if (outidx[u] >= 0) {
#pragma unroll
      for (int q = 0; q < R; ++q) buf[outidx[u] + q * Ns * se] = v[u][q];
    }
  }
  __syncthreads();
}

template <typename T, int DIR, int VPT>
__device__ void fft_lds2(cplx<T>* buf, int nb, int nb2, int sb1, int sb2, int se, const FFTFactors& f,
                         const cplx<T>* tw) {
  int Ns = 1;
  for (int p = 0; p < f.n; ++p) {
    const int R = f.r[p];
    switch (R) {
      case 8: fft_pass2<T, DIR, 8, VPT>(buf, nb, nb2, sb1, sb2, f.L, se, Ns, tw); break;
      case 4: fft_pass2<T, DIR, 4, VPT>(buf, nb, nb2, sb1, sb2, f.L, se, Ns, tw); break;
      case 2: fft_pass2<T, DIR, 2, VPT>(buf, nb, nb2, sb1, sb2, f.L, se, Ns, tw); break;
      case 3: fft_pass2<T, DIR, 3, VPT>(buf, nb, nb2, sb1, sb2, f.L, se, Ns, tw); break;
      case 5: fft_pass2<T, DIR, 5, VPT>(buf, nb, nb2, sb1, sb2, f.L, se, Ns, tw); break;
      default: break;
    }
    Ns *= R;
  }
}

}  // namespace omx
