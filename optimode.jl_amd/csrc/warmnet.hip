// Forward and backward kernels of the warm-start network (layers in warmnet.hpp), FP32, and
// its Adam training loop over a device-resident data set.  Every reduction runs in a fixed
// order (no atomics): training is deterministic for a given seed.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <random>

#include "warmnet.hpp"

namespace omx {

namespace {

constexpr int TB = 256;
constexpr int TS = 16;   // output tile edge of a convolution block (TS x TS = TB threads)
constexpr int COB = 8;   // output channels per convolution block
constexpr int WG = 4;    // output channels per weight-gradient block
constexpr int NBL = 256; // blocks of the field-loss reduction

// 3x3 convolution, zero padding: out[b][o][y][x] = act(bias[o] + sum_i sum_t
// in[b][i][y + t/3 - 1][x + t%3 - 1] w(o, i, t)) for one TS x TS tile, COB output channels and
// one sample per block.  Forward: w(o, i, t) = w[(o cin + i) 9 + t].  BWD (data gradient of a
// forward convolution whose output channels are this call's cin and whose input channels are
// this call's cout): w(o, i, t) = w[(i cout + o) 9 + 8 - t].
template <bool BWD>
__global__ __launch_bounds__(TB) void k_conv(const float* in, int cin, const float* w, const float* bias, float* out,
                                             int cout, int res, int relu) {
  __shared__ float tile[TS + 2][TS + 3];
  __shared__ float ws[COB][9];
  const int ntx = (res + TS - 1) / TS;
  const int ty0 = (int(blockIdx.x) / ntx) * TS, tx0 = (int(blockIdx.x) % ntx) * TS;
  const int o0 = blockIdx.y * COB;
  const int b = blockIdx.z;
  const int ly = threadIdx.x / TS, lx = threadIdx.x % TS;
  const int y = ty0 + ly, x = tx0 + lx;
  const long plane = long(res) * res;
  const float* inb = in + long(b) * cin * plane;
  float acc[COB];
#pragma unroll
  for (int j = 0; j < COB; ++j) acc[j] = 0.f;
  for (int i = 0; i < cin; ++i) {
    const float* ip = inb + long(i) * plane;
    for (int e = threadIdx.x; e < (TS + 2) * (TS + 2); e += TB) {
      const int r = e / (TS + 2), c = e - r * (TS + 2);
      const int yy = ty0 + r - 1, xx = tx0 + c - 1;
      tile[r][c] = (yy >= 0 && yy < res && xx >= 0 && xx < res) ? ip[long(yy) * res + xx] : 0.f;
    }
    if (threadIdx.x < COB * 9) {
      const int j = threadIdx.x / 9, t = threadIdx.x - (threadIdx.x / 9) * 9, o = o0 + j;
      float v = 0.f;
      if (o < cout) v = BWD ? w[(long(i) * cout + o) * 9 + 8 - t] : w[(long(o) * cin + i) * 9 + t];
      ws[j][t] = v;
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < 9; ++t) {
      const float v = tile[ly + t / 3][lx + t % 3];
#pragma unroll
      for (int j = 0; j < COB; ++j) acc[j] += v * ws[j][t];
    }
    __syncthreads();
  }
  if (y >= res || x >= res) return;
#pragma unroll
  for (int j = 0; j < COB; ++j) {
    const int o = o0 + j;
    if (o < cout) {
      float v = acc[j] + (bias ? bias[o] : 0.f);
      if (relu) v = fmaxf(v, 0.f);
      out[(long(b) * cout + o) * plane + long(y) * res + x] = v;
    }
  }
}

// The same convolution on the matrix cores: an implicit GEMM, out (16 output channels x the
// 256 pixels of a TS x TS tile) = W (16 x 9 cin) . im2col(in) (9 cin x 256), contracted in
// steps of 16 with v_mfma_f32_16x16x16bf16_1k: bf16 operands (round to nearest even of the
// FP32 weights and activations), FP32 accumulation.  The block stages the input halo of all
// cin channels and its 16 output channels' weights in LDS (FP32); wave w computes tile rows
// 4w .. 4w + 3, one 16 x 16 MFMA tile (16 channels x the 16 pixels of a row) each.  Operand
// layout of the instruction: A lane l -> (row l%16, k 4(l/16) + j), B lane l -> (k 4(l/16)
// + j, column l%16), accumulator register j of lane l -> (row 4(l/16) + j, column l%16).
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
constexpr int CMX_CO = 16;   // output channels per MFMA convolution block
constexpr int HALO = (TS + 2) * (TS + 2);

__device__ __forceinline__ short bf16_rne(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return short(u >> 16);
}

template <bool BWD>
__global__ __launch_bounds__(TB) void k_conv_mfma(const float* in, int cin, const float* w, const float* bias,
                                                  float* out, int cout, int res, int relu) {
  extern __shared__ __align__(16) float smf[];
  const int kin = cin * 9, kp = ((kin + 15) / 16) * 16;
  float* tile = smf;                 // [cin][TS + 2][TS + 2]
  float* ws = smf + long(cin) * HALO;  // [CMX_CO][kp]
  const int ntx = (res + TS - 1) / TS;
  const int ty0 = (int(blockIdx.x) / ntx) * TS, tx0 = (int(blockIdx.x) % ntx) * TS;
  const int o0 = blockIdx.y * CMX_CO;
  const int b = blockIdx.z;
  const long plane = long(res) * res;
  const float* inb = in + long(b) * cin * plane;
  for (int e = threadIdx.x; e < cin * HALO; e += TB) {
    const int i = e / HALO, q = e - i * HALO, r = q / (TS + 2), c = q - r * (TS + 2);
    const int yy = ty0 + r - 1, xx = tx0 + c - 1;
    tile[e] = (yy >= 0 && yy < res && xx >= 0 && xx < res) ? inb[long(i) * plane + long(yy) * res + xx] : 0.f;
  }
  for (int e = threadIdx.x; e < CMX_CO * kp; e += TB) {
    const int j = e / kp, k = e - j * kp, o = o0 + j;
    float v = 0.f;
    if (o < cout && k < kin) {
      const int i = k / 9, t = k - i * 9;
      v = BWD ? w[(long(i) * cout + o) * 9 + 8 - t] : w[(long(o) * cin + i) * 9 + t];
    }
    ws[e] = v;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int lr = lane & 15, lg = lane >> 4;
  f32x4 acc[4];
#pragma unroll
  for (int rr = 0; rr < 4; ++rr) acc[rr] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < kp; k0 += 16) {
    bf16x4 a;
    int off[4];
    bool ok[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int k = k0 + 4 * lg + j;
      a[j] = bf16_rne(ws[lr * kp + k]);
      ok[j] = k < kin;
      const int i = ok[j] ? k / 9 : 0, t = k - (k / 9) * 9;
      off[j] = i * HALO + (t / 3) * (TS + 2) + lr + t % 3;  // tap t of pixel (row 0, column lr)
    }
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) {
      const int ly = wv * 4 + rr;
      bf16x4 bv;
#pragma unroll
      for (int j = 0; j < 4; ++j) bv[j] = ok[j] ? bf16_rne(tile[off[j] + ly * (TS + 2)]) : short(0);
      acc[rr] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bv, acc[rr], 0, 0, 0);
    }
  }
  const int x = tx0 + lr;
#pragma unroll
  for (int rr = 0; rr < 4; ++rr) {
    const int y = ty0 + wv * 4 + rr;
    if (y >= res || x >= res) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int o = o0 + 4 * lg + j;
      if (o >= cout) continue;
      float v = acc[rr][j] + (bias ? bias[o] : 0.f);
      if (relu) v = fmaxf(v, 0.f);
      out[(long(b) * cout + o) * plane + long(y) * res + x] = v;
    }
  }
}

// Weight gradient: dw[(o cin + i) 9 + t] = sum_b sum_p dout[b][o][p] in[b][i][p + off_t] and
// (blocks with i = 0) db[o] = sum_b sum_p dout[b][o][p]; block (i, group of WG outputs); each
// value reduced per wavefront by a fixed butterfly, then over the wavefronts in order.
__global__ __launch_bounds__(TB) void k_conv_wgrad(const float* in, int cin, const float* dout, int cout, int res,
                                                   int B, float* dw, float* db) {
  const int i = blockIdx.x, o0 = blockIdx.y * WG;
  const long plane = long(res) * res;
  float acc[WG][9], bacc[WG];
#pragma unroll
  for (int j = 0; j < WG; ++j) {
    bacc[j] = 0.f;
#pragma unroll
    for (int t = 0; t < 9; ++t) acc[j][t] = 0.f;
  }
  for (long e = threadIdx.x; e < long(B) * plane; e += TB) {
    const int b = int(e / plane);
    const long p = e - long(b) * plane;
    const int y = int(p / res), x = int(p - long(p / res) * res);
    const float* ip = in + (long(b) * cin + i) * plane;
    float v[9];
#pragma unroll
    for (int t = 0; t < 9; ++t) {
      const int yy = y + t / 3 - 1, xx = x + t % 3 - 1;
      v[t] = (yy >= 0 && yy < res && xx >= 0 && xx < res) ? ip[long(yy) * res + xx] : 0.f;
    }
#pragma unroll
    for (int j = 0; j < WG; ++j) {
      const int o = o0 + j;
      const float d = o < cout ? dout[(long(b) * cout + o) * plane + p] : 0.f;
      bacc[j] += d;
#pragma unroll
      for (int t = 0; t < 9; ++t) acc[j][t] += d * v[t];
    }
  }
  constexpr int NV = WG * 10, NW = TB / 64;
  __shared__ float red[NW][NV];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    float v = q < WG * 9 ? acc[q / 9][q % 9] : bacc[q - WG * 9];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) red[wave][q] = v;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    const int q = threadIdx.x;
    float s = 0.f;
#pragma unroll
    for (int w = 0; w < NW; ++w) s += red[w][q];
    if (q < WG * 9) {
      const int o = o0 + q / 9;
      if (o < cout) dw[(long(o) * cin + i) * 9 + q % 9] = s;
    } else {
      const int o = o0 + (q - WG * 9);
      if (i == 0 && o < cout) db[o] = s;
    }
  }
}

// 2x2 average pooling: out (B C planes of h x h) from in (planes of 2h x 2h)
__global__ __launch_bounds__(TB) void k_pool(const float* in, float* out, long nout, int h) {
  for (long e = long(blockIdx.x) * TB + threadIdx.x; e < nout; e += long(gridDim.x) * TB) {
    const long hh = long(h) * h;
    const long bc = e / hh, p = e - bc * hh;
    const int y = int(p / h), x = int(p - long(p / h) * h);
    const int R = 2 * h;
    const float* ip = in + bc * 4 * hh + long(2 * y) * R + 2 * x;
    out[e] = 0.25f * (ip[0] + ip[1] + ip[R] + ip[R + 1]);
  }
}

// din (planes of R x R) (+)= dout / 4 at (Y/2, X/2)
__global__ __launch_bounds__(TB) void k_pool_bwd(const float* dout, float* din, long nin, int R, int accumulate) {
  for (long e = long(blockIdx.x) * TB + threadIdx.x; e < nin; e += long(gridDim.x) * TB) {
    const long RR = long(R) * R;
    const int h = R / 2;
    const long bc = e / RR, p = e - bc * RR;
    const int Y = int(p / R), X = int(p - long(p / R) * R);
    const float v = 0.25f * dout[bc * long(h) * h + long(Y / 2) * h + X / 2];
    din[e] = accumulate ? din[e] + v : v;
  }
}

// nearest-neighbour upsampling into channels c0.. of dst: dst[b][c0 + c][Y][X] = src[b][c][Y/2][X/2]
__global__ __launch_bounds__(TB) void k_up(const float* src, int C, int h, float* dst, int cdst, int c0, int B) {
  const int R = 2 * h;
  const long RR = long(R) * R, n = long(B) * C * RR;
  for (long e = long(blockIdx.x) * TB + threadIdx.x; e < n; e += long(gridDim.x) * TB) {
    const long bc = e / RR, p = e - bc * RR;
    const int b = int(bc / C), c = int(bc - long(bc / C) * C);
    const int Y = int(p / R), X = int(p - long(p / R) * R);
    dst[(long(b) * cdst + c0 + c) * RR + p] = src[bc * long(h) * h + long(Y / 2) * h + X / 2];
  }
}

// dsrc[b][c][y][x] = sum of the 2x2 block of ddst[b][c0 + c]
__global__ __launch_bounds__(TB) void k_up_bwd(const float* ddst, int cdst, int c0, float* dsrc, int C, int h, int B) {
  const long hh = long(h) * h, n = long(B) * C * hh;
  const int R = 2 * h;
  for (long e = long(blockIdx.x) * TB + threadIdx.x; e < n; e += long(gridDim.x) * TB) {
    const long bc = e / hh, p = e - bc * hh;
    const int b = int(bc / C), c = int(bc - long(bc / C) * C);
    const int y = int(p / h), x = int(p - long(p / h) * h);
    const float* dp = ddst + (long(b) * cdst + c0 + c) * 4 * hh + long(2 * y) * R + 2 * x;
    dsrc[e] = dp[0] + dp[1] + dp[R] + dp[R + 1];
  }
}

// dst[b][c0 + c][p] = src[b][c][p]
__global__ __launch_bounds__(TB) void k_put_ch(const float* src, int C, long plane, float* dst, int cdst, int c0,
                                               int B) {
  const long n = long(B) * C * plane;
  for (long e = long(blockIdx.x) * TB + threadIdx.x; e < n; e += long(gridDim.x) * TB) {
    const long bc = e / plane, p = e - bc * plane;
    const int b = int(bc / C), c = int(bc - long(bc / C) * C);
    dst[(long(b) * cdst + c0 + c) * plane + p] = src[e];
  }
}

// dst[b][c][p] = src[b][c0 + c][p]
__global__ __launch_bounds__(TB) void k_take_ch(const float* src, int csrc, int c0, float* dst, int C, long plane,
                                                int B) {
  const long n = long(B) * C * plane;
  for (long e = long(blockIdx.x) * TB + threadIdx.x; e < n; e += long(gridDim.x) * TB) {
    const long bc = e / plane, p = e - bc * plane;
    const int b = int(bc / C), c = int(bc - long(bc / C) * C);
    dst[e] = src[(long(b) * csrc + c0 + c) * plane + p];
  }
}

__global__ __launch_bounds__(TB) void k_relu_bwd(float* d, const float* a, long n) {
  for (long e = long(blockIdx.x) * TB + threadIdx.x; e < n; e += long(gridDim.x) * TB)
    if (!(a[e] > 0.f)) d[e] = 0.f;
}

// g[bc] = mean over the plane of a[bc]; one block per (b, c)
__global__ __launch_bounds__(TB) void k_gap(const float* a, long plane, float* g) {
  const float* ap = a + long(blockIdx.x) * plane;
  float s = 0.f;
  for (long p = threadIdx.x; p < plane; p += TB) s += ap[p];
  __shared__ float red[TB / 64];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    float t = 0.f;
    for (int w = 0; w < TB / 64; ++w) t += red[w];
    g[blockIdx.x] = t / float(plane);
  }
}

// k[b][j] = bias[j] + sum_c W[j nf + c] g[b][c]
__global__ __launch_bounds__(TB) void k_dense(const float* g, const float* W, const float* bias, float* k, int B,
                                              int nout, int nf) {
  for (int e = blockIdx.x * TB + threadIdx.x; e < B * nout; e += gridDim.x * TB) {
    const int b = e / nout, j = e - (e / nout) * nout;
    float s = bias[j];
    for (int c = 0; c < nf; ++c) s += W[j * nf + c] * g[b * nf + c];
    k[e] = s;
  }
}

// dW[j][c] = sum_b dk[b][j] g[b][c], dbias[j] = sum_b dk[b][j], dg[b][c] = sum_j W[j][c] dk[b][j]
__global__ __launch_bounds__(TB) void k_dense_bwd(const float* dk, const float* g, const float* W, float* dW,
                                                  float* dbias, float* dg, int B, int nout, int nf) {
  for (int e = threadIdx.x; e < nout * nf; e += TB) {
    const int j = e / nf, c = e - (e / nf) * nf;
    float s = 0.f;
    for (int b = 0; b < B; ++b) s += dk[b * nout + j] * g[b * nf + c];
    dW[e] = s;
  }
  for (int j = threadIdx.x; j < nout; j += TB) {
    float s = 0.f;
    for (int b = 0; b < B; ++b) s += dk[b * nout + j];
    dbias[j] = s;
  }
  for (int e = threadIdx.x; e < B * nf; e += TB) {
    const int b = e / nf, c = e - (e / nf) * nf;
    float s = 0.f;
    for (int j = 0; j < nout; ++j) s += W[j * nf + c] * dk[b * nout + j];
    dg[e] = s;
  }
}

// da[bc][p] += dg[bc] / plane
__global__ __launch_bounds__(TB) void k_gap_bwd(float* da, const float* dg, long plane, long n) {
  for (long e = long(blockIdx.x) * TB + threadIdx.x; e < n; e += long(gridDim.x) * TB)
    da[e] += dg[e / plane] / float(plane);
}

// d = s (y - t); part[blockIdx.x] = sum over this block's entries of (y - t)^2 (FP64)
__global__ __launch_bounds__(TB) void k_mse(const float* y, const float* t, float* d, long n, float s, double* part) {
  double acc = 0.0;
  for (long e = long(blockIdx.x) * TB + threadIdx.x; e < n; e += long(gridDim.x) * TB) {
    const float r = y[e] - t[e];
    d[e] = s * r;
    acc += double(r) * double(r);
  }
  __shared__ double red[TB / 64];
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double u = 0.0;
    for (int w = 0; w < TB / 64; ++w) u += red[w];
    part[blockIdx.x] = u;
  }
}

// loss[slot] += (sum part[0..nbf)) / nf + wk part[nbf] / nk
__global__ void k_loss_sum(const double* part, int nbf, double nf, double wk, double nk, double* loss, int slot) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double s = 0.0;
  for (int b = 0; b < nbf; ++b) s += part[b];
  loss[slot] += s / nf + wk * part[nbf] / nk;
}

__global__ __launch_bounds__(TB) void k_adam(float* p, const float* g, float* m, float* v, long n, float lr, float b1,
                                             float b2, float eps, float c1, float c2) {
  for (long e = long(blockIdx.x) * TB + threadIdx.x; e < n; e += long(gridDim.x) * TB) {
    const float ge = g[e];
    const float me = b1 * m[e] + (1.f - b1) * ge;
    const float ve = b2 * v[e] + (1.f - b2) * ge * ge;
    m[e] = me;
    v[e] = ve;
    p[e] -= lr * (me / c1) / (sqrtf(ve / c2) + eps);
  }
}

// dst[j][0..per) = src[idx[j]][0..per) for j < B
__global__ __launch_bounds__(TB) void k_gather(const float* src, const int* idx, long per, float* dst, int B) {
  const long n = long(B) * per;
  for (long e = long(blockIdx.x) * TB + threadIdx.x; e < n; e += long(gridDim.x) * TB) {
    const long j = e / per, q = e - j * per;
    dst[e] = src[long(idx[j]) * per + q];
  }
}

// x[(c hw + bx) hw + by] = normalised mean of eps6 component c over pixel block (bx, by)
// (fx x fy pixels), c < 6; channel 6 = normalised omega
__global__ __launch_bounds__(TB) void k_ws_features(const double* eps6, int nx, int ny, int hw, double omega,
                                                    float* x) {
  const int fx = nx / hw, fy = ny / hw;
  const long n = 7L * hw * hw;
  for (long e = long(blockIdx.x) * TB + threadIdx.x; e < n; e += long(gridDim.x) * TB) {
    const int c = int(e / (long(hw) * hw));
    const int q = int(e - long(c) * hw * hw);
    const int bx = q / hw, by = q - (q / hw) * hw;
    double v;
    if (c == 6) {
      v = (omega - WS_OMEGA0) / WS_OMEGA_SCALE;
    } else {
      double sum = 0.0;
      for (int iy = by * fy; iy < (by + 1) * fy; ++iy)
        for (int ix = bx * fx; ix < (bx + 1) * fx; ++ix) sum += eps6[6 * (long(ix) + long(nx) * iy) + c];
      const double m = sum / double(fx * fy);
      v = c < 3 ? (m - WS_DIAG0) / WS_DIAG_SCALE : m / WS_OFF_SCALE;
    }
    x[e] = float(v);
  }
}

// inverse DFT along the second map axis: t[(jc hw + x) hw + gy] = sum_y u_jc[x][y] e^{+2 pi i gy y / hw}
// with u_jc = y[4j + 2c] + i y[4j + 2c + 1] (jc = 2j + c)
__global__ __launch_bounds__(TB) void k_ws_idft1(const float* yf, int nev, int hw, c64* t) {
  const long hh = long(hw) * hw, n = 2L * nev * hh;
  for (long e = long(blockIdx.x) * TB + threadIdx.x; e < n; e += long(gridDim.x) * TB) {
    const int jc = int(e / hh);
    const int q = int(e - long(jc) * hh);
    const int x = q / hw, gy = q - (q / hw) * hw;
    const float* re = yf + long(2 * jc) * hh + long(x) * hw;
    const float* im = re + hh;
    double sr = 0.0, si = 0.0;
    for (int y = 0; y < hw; ++y) {
      double sn, cs;
      sincospi(2.0 * double((gy * y) % hw) / double(hw), &sn, &cs);
      sr += re[y] * cs - im[y] * sn;
      si += re[y] * sn + im[y] * cs;
    }
    t[e] = c64{sr, si};
  }
}

// inverse DFT along the first axis with the 1/hw^2 of ifft2: band j's vector entry
// c + 2(gx + hw gy) = sum_x t_jc[x][gy] e^{+2 pi i gx x / hw} / hw^2
__global__ __launch_bounds__(TB) void k_ws_idft2(const c64* t, int nev, int hw, c64* out) {
  const long hh = long(hw) * hw, n = 2L * nev * hh;
  const double sc = 1.0 / double(hh);
  for (long e = long(blockIdx.x) * TB + threadIdx.x; e < n; e += long(gridDim.x) * TB) {
    const int jc = int(e / hh);
    const int q = int(e - long(jc) * hh);
    const int gx = q / hw, gy = q - (q / hw) * hw;
    const c64* tc = t + long(jc) * hh + gy;
    double sr = 0.0, si = 0.0;
    for (int x = 0; x < hw; ++x) {
      double sn, cs;
      sincospi(2.0 * double((gx * x) % hw) / double(hw), &sn, &cs);
      const c64 v = tc[long(x) * hw];
      sr += v.x * cs - v.y * sn;
      si += v.x * sn + v.y * cs;
    }
    const int j = jc / 2, c = jc - 2 * (jc / 2);
    out[long(j) * 2 * hh + c + 2 * (long(gx) + long(hw) * gy)] = c64{sr * sc, si * sc};
  }
}

// unit norm per band (one block per band, fixed-order reduction)
__global__ __launch_bounds__(TB) void k_ws_normalise(c64* v, long per) {
  c64* b = v + long(blockIdx.x) * per;
  double s = 0.0;
  for (long e = threadIdx.x; e < per; e += TB) s += b[e].x * b[e].x + b[e].y * b[e].y;
  __shared__ double red[TB / 64];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  double t = 0.0;
  for (int w = 0; w < TB / 64; ++w) t += red[w];
  const double f = t > 0.0 ? 1.0 / sqrt(t) : 0.0;
  for (long e = threadIdx.x; e < per; e += TB) b[e] = c64{b[e].x * f, b[e].y * f};
}

unsigned grid_for(long n) { return unsigned(std::max<long>(1, std::min<long>((n + TB - 1) / TB, 4096))); }

}  // namespace

WarmNet::WarmNet(int device, int hw, int nev, uint64_t seed) : device_(device), hw_(hw), nev_(nev) {
  if (hw < 4 || hw > 256 || hw % 4 != 0) throw Error(OMX_ERR_BAD_SHAPE, "warm-start map size must be a multiple of 4 in 4..256");
  if (nev < 1 || nev > 16) throw Error(OMX_ERR_UNSUPPORTED, "nev must be in 1..16");
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count)
    throw Error(OMX_ERR_NO_DEVICE, "no such HIP device");
  OMX_HIP_CHECK(hipSetDevice(device));
  OMX_HIP_CHECK(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking));
  const int dims[6][3] = {{NIN, 16, hw}, {16, 32, hw / 2}, {32, 32, hw / 4},
                          {64, 32, hw / 2}, {48, 32, hw}, {32, 4 * nev, hw}};
  long off = 0;
  for (int l = 0; l < 6; ++l) {
    c_[l] = Conv{dims[l][0], dims[l][1], dims[l][2], off, off + long(dims[l][0]) * dims[l][1] * 9};
    off = c_[l].b + dims[l][1];
  }
  dw_ = off;
  db_ = dw_ + long(nev) * 32;
  nparams_ = db_ + nev;
  // He-normal weights (output layers scaled down), zero biases
  std::vector<float> h(nparams_, 0.f);
  std::mt19937_64 rng(seed);
  std::normal_distribution<double> nd(0.0, 1.0);
  for (int l = 0; l < 6; ++l) {
    const double sd = std::sqrt(2.0 / (9.0 * c_[l].ci)) * (l == 5 ? 0.1 : 1.0);
    for (long q = c_[l].w; q < c_[l].b; ++q) h[q] = float(sd * nd(rng));
  }
  for (long q = dw_; q < db_; ++q) h[q] = float(0.1 * std::sqrt(1.0 / 32.0) * nd(rng));
  P_.alloc(sizeof(float) * nparams_);
  G_.alloc(sizeof(float) * nparams_);
  M1_.alloc(sizeof(float) * nparams_);
  M2_.alloc(sizeof(float) * nparams_);
  set_params(h.data());
}

WarmNet::~WarmNet() {
  if (s_) {
    (void)hipSetDevice(device_);
    (void)hipStreamSynchronize(s_);
    (void)hipStreamDestroy(s_);
  }
}

void WarmNet::get_params(float* out) {
  OMX_HIP_CHECK(hipSetDevice(device_));
  OMX_HIP_CHECK(hipMemcpyAsync(out, P_.as<void>(), sizeof(float) * nparams_, hipMemcpyDeviceToHost, s_));
  OMX_HIP_CHECK(hipStreamSynchronize(s_));
}

void WarmNet::set_params(const float* in) {
  OMX_HIP_CHECK(hipSetDevice(device_));
  OMX_HIP_CHECK(hipMemcpyAsync(P_.as<void>(), in, sizeof(float) * nparams_, hipMemcpyHostToDevice, s_));
  OMX_HIP_CHECK(hipStreamSynchronize(s_));
}

void WarmNet::ensure_batch(int B) {
  if (B <= B_) return;
  const long H2 = long(hw_) * hw_, q2 = H2 / 4, r2 = H2 / 16, of = 4L * nev_;
  const size_t f = sizeof(float) * size_t(B);
  x_.alloc(f * NIN * H2);
  a1_.alloc(f * 16 * H2);
  p1_.alloc(f * 16 * q2);
  a2_.alloc(f * 32 * q2);
  p2_.alloc(f * 32 * r2);
  a3_.alloc(f * 32 * r2);
  g_.alloc(f * 32);
  k_.alloc(f * nev_);
  cat4_.alloc(f * 64 * q2);
  a4_.alloc(f * 32 * q2);
  cat5_.alloc(f * 48 * H2);
  a5_.alloc(f * 32 * H2);
  y_.alloc(f * of * H2);
  dy_.alloc(f * of * H2);
  da5_.alloc(f * 32 * H2);
  dcat5_.alloc(f * 48 * H2);
  da4_.alloc(f * 32 * q2);
  dcat4_.alloc(f * 64 * q2);
  da3_.alloc(f * 32 * r2);
  dp2_.alloc(f * 32 * r2);
  da2_.alloc(f * 32 * q2);
  dp1_.alloc(f * 16 * q2);
  da1_.alloc(f * 16 * H2);
  dk_.alloc(f * nev_);
  dg_.alloc(f * 32);
  part_.alloc(sizeof(double) * (NBL + 1));
  B_ = B;
}

static size_t conv_mfma_lds(int cin) {
  return sizeof(float) * (size_t(cin) * HALO + size_t(CMX_CO) * (((cin * 9 + 15) / 16) * 16));
}

void WarmNet::conv(const Conv& c, const float* in, float* out, int B, bool relu) {
  const int nt = (c.res + TS - 1) / TS;
  const float* P = P_.as<float>();
  if (bf16_)
    hipLaunchKernelGGL(k_conv_mfma<false>, dim3(nt * nt, (c.co + CMX_CO - 1) / CMX_CO, B), dim3(TB),
                       conv_mfma_lds(c.ci), s_, in, c.ci, P + c.w, P + c.b, out, c.co, c.res, relu ? 1 : 0);
  else
    hipLaunchKernelGGL(k_conv<false>, dim3(nt * nt, (c.co + COB - 1) / COB, B), dim3(TB), 0, s_, in, c.ci, P + c.w,
                       P + c.b, out, c.co, c.res, relu ? 1 : 0);
}

void WarmNet::conv_bwd(const Conv& c, const float* in, const float* dout, float* din, int B) {
  float* G = G_.as<float>();
  hipLaunchKernelGGL(k_conv_wgrad, dim3(c.ci, (c.co + WG - 1) / WG), dim3(TB), 0, s_, in, c.ci, dout, c.co, c.res, B,
                     G + c.w, G + c.b);
  if (din) {
    const int nt = (c.res + TS - 1) / TS;
    if (bf16_)
      hipLaunchKernelGGL(k_conv_mfma<true>, dim3(nt * nt, (c.ci + CMX_CO - 1) / CMX_CO, B), dim3(TB),
                         conv_mfma_lds(c.co), s_, dout, c.co, P_.as<float>() + c.w, nullptr, din, c.ci, c.res, 0);
    else
      hipLaunchKernelGGL(k_conv<true>, dim3(nt * nt, (c.ci + COB - 1) / COB, B), dim3(TB), 0, s_, dout, c.co,
                         P_.as<float>() + c.w, nullptr, din, c.ci, c.res, 0);
  }
}

void WarmNet::fwd(int B) {
  const int H = hw_, q = H / 2, r = H / 4;
  const long H2 = long(H) * H, q2 = long(q) * q, r2 = long(r) * r;
  const float* P = P_.as<float>();
  conv(c_[0], x_.as<float>(), a1_.as<float>(), B, true);
  hipLaunchKernelGGL(k_pool, dim3(grid_for(long(B) * 16 * q2)), dim3(TB), 0, s_, a1_.as<float>(), p1_.as<float>(),
                     long(B) * 16 * q2, q);
  conv(c_[1], p1_.as<float>(), a2_.as<float>(), B, true);
  hipLaunchKernelGGL(k_pool, dim3(grid_for(long(B) * 32 * r2)), dim3(TB), 0, s_, a2_.as<float>(), p2_.as<float>(),
                     long(B) * 32 * r2, r);
  conv(c_[2], p2_.as<float>(), a3_.as<float>(), B, true);
  hipLaunchKernelGGL(k_gap, dim3(B * 32), dim3(TB), 0, s_, a3_.as<float>(), r2, g_.as<float>());
  hipLaunchKernelGGL(k_dense, dim3(grid_for(long(B) * nev_)), dim3(TB), 0, s_, g_.as<float>(), P + dw_, P + db_,
                     k_.as<float>(), B, nev_, 32);
  hipLaunchKernelGGL(k_up, dim3(grid_for(long(B) * 32 * q2)), dim3(TB), 0, s_, a3_.as<float>(), 32, r,
                     cat4_.as<float>(), 64, 0, B);
  hipLaunchKernelGGL(k_put_ch, dim3(grid_for(long(B) * 32 * q2)), dim3(TB), 0, s_, a2_.as<float>(), 32, q2,
                     cat4_.as<float>(), 64, 32, B);
  conv(c_[3], cat4_.as<float>(), a4_.as<float>(), B, true);
  hipLaunchKernelGGL(k_up, dim3(grid_for(long(B) * 32 * H2)), dim3(TB), 0, s_, a4_.as<float>(), 32, q,
                     cat5_.as<float>(), 48, 0, B);
  hipLaunchKernelGGL(k_put_ch, dim3(grid_for(long(B) * 16 * H2)), dim3(TB), 0, s_, a1_.as<float>(), 16, H2,
                     cat5_.as<float>(), 48, 32, B);
  conv(c_[4], cat5_.as<float>(), a5_.as<float>(), B, true);
  conv(c_[5], a5_.as<float>(), y_.as<float>(), B, false);
  OMX_HIP_CHECK(hipGetLastError());
}

void WarmNet::bwd(int B, const float* tf, const float* tk, float wk, int slot) {
  const int H = hw_, q = H / 2, r = H / 4;
  const long H2 = long(H) * H, q2 = long(q) * q, r2 = long(r) * r, of = 4L * nev_;
  const long nf = long(B) * of * H2, nk = long(B) * nev_;
  const float* P = P_.as<float>();
  float* G = G_.as<float>();
  double* part = part_.as<double>();
  hipLaunchKernelGGL(k_mse, dim3(NBL), dim3(TB), 0, s_, y_.as<float>(), tf, dy_.as<float>(), nf, float(2.0 / double(nf)),
                     part);
  hipLaunchKernelGGL(k_mse, dim3(1), dim3(TB), 0, s_, k_.as<float>(), tk, dk_.as<float>(), nk,
                     float(2.0 * wk / double(nk)), part + NBL);
  hipLaunchKernelGGL(k_loss_sum, dim3(1), dim3(64), 0, s_, part, NBL, double(nf), double(wk), double(nk),
                     lossd_.as<double>(), slot);
  // c6, c5 (input [up(c4) | c1])
  conv_bwd(c_[5], a5_.as<float>(), dy_.as<float>(), da5_.as<float>(), B);
  hipLaunchKernelGGL(k_relu_bwd, dim3(grid_for(long(B) * 32 * H2)), dim3(TB), 0, s_, da5_.as<float>(), a5_.as<float>(),
                     long(B) * 32 * H2);
  conv_bwd(c_[4], cat5_.as<float>(), da5_.as<float>(), dcat5_.as<float>(), B);
  hipLaunchKernelGGL(k_up_bwd, dim3(grid_for(long(B) * 32 * q2)), dim3(TB), 0, s_, dcat5_.as<float>(), 48, 0,
                     da4_.as<float>(), 32, q, B);
  hipLaunchKernelGGL(k_take_ch, dim3(grid_for(long(B) * 16 * H2)), dim3(TB), 0, s_, dcat5_.as<float>(), 48, 32,
                     da1_.as<float>(), 16, H2, B);
  // c4 (input [up(c3) | c2])
  hipLaunchKernelGGL(k_relu_bwd, dim3(grid_for(long(B) * 32 * q2)), dim3(TB), 0, s_, da4_.as<float>(), a4_.as<float>(),
                     long(B) * 32 * q2);
  conv_bwd(c_[3], cat4_.as<float>(), da4_.as<float>(), dcat4_.as<float>(), B);
  hipLaunchKernelGGL(k_up_bwd, dim3(grid_for(long(B) * 32 * r2)), dim3(TB), 0, s_, dcat4_.as<float>(), 64, 0,
                     da3_.as<float>(), 32, r, B);
  hipLaunchKernelGGL(k_take_ch, dim3(grid_for(long(B) * 32 * q2)), dim3(TB), 0, s_, dcat4_.as<float>(), 64, 32,
                     da2_.as<float>(), 32, q2, B);
  // head: dense and the pixel mean of c3
  hipLaunchKernelGGL(k_dense_bwd, dim3(1), dim3(TB), 0, s_, dk_.as<float>(), g_.as<float>(), P + dw_, G + dw_, G + db_,
                     dg_.as<float>(), B, nev_, 32);
  hipLaunchKernelGGL(k_gap_bwd, dim3(grid_for(long(B) * 32 * r2)), dim3(TB), 0, s_, da3_.as<float>(), dg_.as<float>(),
                     r2, long(B) * 32 * r2);
  // c3, pool, c2, pool, c1
  hipLaunchKernelGGL(k_relu_bwd, dim3(grid_for(long(B) * 32 * r2)), dim3(TB), 0, s_, da3_.as<float>(), a3_.as<float>(),
                     long(B) * 32 * r2);
  conv_bwd(c_[2], p2_.as<float>(), da3_.as<float>(), dp2_.as<float>(), B);
  hipLaunchKernelGGL(k_pool_bwd, dim3(grid_for(long(B) * 32 * q2)), dim3(TB), 0, s_, dp2_.as<float>(), da2_.as<float>(),
                     long(B) * 32 * q2, q, 1);
  hipLaunchKernelGGL(k_relu_bwd, dim3(grid_for(long(B) * 32 * q2)), dim3(TB), 0, s_, da2_.as<float>(), a2_.as<float>(),
                     long(B) * 32 * q2);
  conv_bwd(c_[1], p1_.as<float>(), da2_.as<float>(), dp1_.as<float>(), B);
  hipLaunchKernelGGL(k_pool_bwd, dim3(grid_for(long(B) * 16 * H2)), dim3(TB), 0, s_, dp1_.as<float>(), da1_.as<float>(),
                     long(B) * 16 * H2, H, 1);
  hipLaunchKernelGGL(k_relu_bwd, dim3(grid_for(long(B) * 16 * H2)), dim3(TB), 0, s_, da1_.as<float>(), a1_.as<float>(),
                     long(B) * 16 * H2);
  conv_bwd(c_[0], x_.as<float>(), da1_.as<float>(), nullptr, B);
  OMX_HIP_CHECK(hipGetLastError());
}

void WarmNet::forward(int B, const float* x, float* fields, float* kout) {
  if (B < 1 || !x) throw Error(OMX_ERR_INVALID_ARG, "bad warm-start forward arguments");
  OMX_HIP_CHECK(hipSetDevice(device_));
  ensure_batch(B);
  const long H2 = long(hw_) * hw_;
  OMX_HIP_CHECK(hipMemcpyAsync(x_.as<void>(), x, sizeof(float) * B * NIN * H2, hipMemcpyHostToDevice, s_));
  fwd(B);
  if (fields)
    OMX_HIP_CHECK(hipMemcpyAsync(fields, y_.as<void>(), sizeof(float) * B * 4 * nev_ * H2, hipMemcpyDeviceToHost, s_));
  if (kout) OMX_HIP_CHECK(hipMemcpyAsync(kout, k_.as<void>(), sizeof(float) * B * nev_, hipMemcpyDeviceToHost, s_));
  OMX_HIP_CHECK(hipStreamSynchronize(s_));
}

double WarmNet::loss_grad(int B, const float* x, const float* tf, const float* tk, float wk, float* grad) {
  if (B < 1 || !x || !tf || !tk) throw Error(OMX_ERR_INVALID_ARG, "bad warm-start loss arguments");
  OMX_HIP_CHECK(hipSetDevice(device_));
  ensure_batch(B);
  const long H2 = long(hw_) * hw_, nf = long(B) * 4 * nev_ * H2;
  DevMem dtf(sizeof(float) * nf), dtk(sizeof(float) * B * nev_);
  lossd_.ensure(sizeof(double));
  OMX_HIP_CHECK(hipMemsetAsync(lossd_.as<void>(), 0, sizeof(double), s_));
  OMX_HIP_CHECK(hipMemcpyAsync(x_.as<void>(), x, sizeof(float) * B * NIN * H2, hipMemcpyHostToDevice, s_));
  OMX_HIP_CHECK(hipMemcpyAsync(dtf.as<void>(), tf, sizeof(float) * nf, hipMemcpyHostToDevice, s_));
  OMX_HIP_CHECK(hipMemcpyAsync(dtk.as<void>(), tk, sizeof(float) * B * nev_, hipMemcpyHostToDevice, s_));
  fwd(B);
  bwd(B, dtf.as<float>(), dtk.as<float>(), wk, 0);
  double loss = 0;
  OMX_HIP_CHECK(hipMemcpyAsync(&loss, lossd_.as<void>(), sizeof(double), hipMemcpyDeviceToHost, s_));
  if (grad) OMX_HIP_CHECK(hipMemcpyAsync(grad, G_.as<void>(), sizeof(float) * nparams_, hipMemcpyDeviceToHost, s_));
  OMX_HIP_CHECK(hipStreamSynchronize(s_));
  return loss;
}

const c64* WarmNet::guess_device(const double* eps6, int nx, int ny, double omega, float* k_host) {
  if (nx % hw_ != 0 || ny % hw_ != 0)
    throw Error(OMX_ERR_BAD_SHAPE, "the grid is not a multiple of the warm-start map size");
  OMX_HIP_CHECK(hipSetDevice(device_));
  ensure_batch(1);
  const long n7 = 7L * hw_ * hw_, per = 2L * hw_ * hw_, nt = long(nev_) * per;
  hipLaunchKernelGGL(k_ws_features, dim3(grid_for(n7)), dim3(TB), 0, s_, eps6, nx, ny, hw_, omega, x_.as<float>());
  fwd(1);
  dft_.ensure(sizeof(c64) * size_t(nt));
  low_.ensure(sizeof(c64) * size_t(nt));
  c64* lowres = low_.as<c64>();
  hipLaunchKernelGGL(k_ws_idft1, dim3(grid_for(nt)), dim3(TB), 0, s_, y_.as<float>(), nev_, hw_, dft_.as<c64>());
  hipLaunchKernelGGL(k_ws_idft2, dim3(grid_for(nt)), dim3(TB), 0, s_, dft_.as<c64>(), nev_, hw_, lowres);
  hipLaunchKernelGGL(k_ws_normalise, dim3(nev_), dim3(TB), 0, s_, lowres, per);
  OMX_HIP_CHECK(hipGetLastError());
  OMX_HIP_CHECK(hipMemcpyAsync(k_host, k_.as<void>(), sizeof(float) * nev_, hipMemcpyDeviceToHost, s_));
  OMX_HIP_CHECK(hipStreamSynchronize(s_));
  return lowres;
}

void solve_k_learned(Solver& S, WarmNet& net, double omega, int nev, double k0_scale, double k_tol, double eig_tol,
                     int max_eigsolves, int maxiter, double* kmags, double* evecs, omx_info* info) {
  S.require_2d("the learned start");
  if (net.device() != S.device) throw Error(OMX_ERR_INVALID_ARG, "the network and the solver are on different devices");
  if (nev < 1 || nev > net.nev()) throw Error(OMX_ERR_INVALID_ARG, "nev must be in 1..(the network's band count)");
  if (!(k0_scale > 0) || !std::isfinite(k0_scale) || !(omega > 0) || !std::isfinite(omega))
    throw Error(OMX_ERR_INVALID_ARG, "omega and k0_scale must be positive and finite");
  std::vector<float> kh(net.nev());
  const c64* low = net.guess_device(S.p64.plan.eps6, S.nx, S.ny, omega, kh.data());
  const double k0 = std::max(1e-3, k0_scale * omega * (WS_N_SHIFT + WS_N_SCALE * double(kh[0])));
  struct Restore {
    Solver& s;
    int hw;
    ~Restore() {
      s.opt.guess_lowres = hw;
      s.guess_dev = nullptr;
    }
  } restore{S, S.opt.guess_lowres};
  S.opt.guess_lowres = net.hw();
  S.guess_dev = low;
  // Hguess is only tested for presence: the vectors come from guess_dev
  S.solve_k(omega, nev, k_tol, eig_tol, max_eigsolves, maxiter, &k0, reinterpret_cast<const double*>(low), nev, kmags,
            evecs, info);
}

void WarmNet::train(int n, const float* x, const float* tf, const float* tk, float wk, int epochs, int batch,
                    float lr, uint64_t seed, double* hist, bool resume) {
  if (n < 1 || epochs < 1 || batch < 1 || !x || !tf || !tk || !(lr > 0))
    throw Error(OMX_ERR_INVALID_ARG, "bad warm-start training arguments");
  OMX_HIP_CHECK(hipSetDevice(device_));
  const int bs = std::min(batch, n);
  const int nb = n / bs;  // minibatches per epoch (the remainder of a shuffle is left out)
  ensure_batch(bs);
  const long H2 = long(hw_) * hw_, perx = NIN * H2, perf = 4L * nev_ * H2;
  DevMem dX(sizeof(float) * n * perx), dF(sizeof(float) * n * perf), dK(sizeof(float) * n * nev_);
  DevMem bF(sizeof(float) * bs * perf), bK(sizeof(float) * bs * nev_), dIdx(sizeof(int) * size_t(epochs) * n);
  OMX_HIP_CHECK(hipMemcpyAsync(dX.as<void>(), x, sizeof(float) * n * perx, hipMemcpyHostToDevice, s_));
  OMX_HIP_CHECK(hipMemcpyAsync(dF.as<void>(), tf, sizeof(float) * n * perf, hipMemcpyHostToDevice, s_));
  OMX_HIP_CHECK(hipMemcpyAsync(dK.as<void>(), tk, sizeof(float) * n * nev_, hipMemcpyHostToDevice, s_));
  // every epoch's shuffle, uploaded once
  std::vector<int> perm(size_t(epochs) * n);
  std::mt19937_64 rng(seed);
  for (int e = 0; e < epochs; ++e) {
    int* pe = perm.data() + size_t(e) * n;
    std::iota(pe, pe + n, 0);
    std::shuffle(pe, pe + n, rng);
  }
  OMX_HIP_CHECK(hipMemcpyAsync(dIdx.as<void>(), perm.data(), sizeof(int) * perm.size(), hipMemcpyHostToDevice, s_));
  lossd_.ensure(sizeof(double) * epochs);
  OMX_HIP_CHECK(hipMemsetAsync(lossd_.as<void>(), 0, sizeof(double) * epochs, s_));
  if (!resume || adam_t_ == 0) {
    OMX_HIP_CHECK(hipMemsetAsync(M1_.as<void>(), 0, sizeof(float) * nparams_, s_));
    OMX_HIP_CHECK(hipMemsetAsync(M2_.as<void>(), 0, sizeof(float) * nparams_, s_));
    adam_t_ = 0;
  }
  const float b1 = 0.9f, b2 = 0.999f;
  for (int e = 0; e < epochs; ++e) {
    for (int m = 0; m < nb; ++m) {
      const int* idx = dIdx.as<int>() + size_t(e) * n + size_t(m) * bs;
      hipLaunchKernelGGL(k_gather, dim3(grid_for(long(bs) * perx)), dim3(TB), 0, s_, dX.as<float>(), idx, perx,
                         x_.as<float>(), bs);
      hipLaunchKernelGGL(k_gather, dim3(grid_for(long(bs) * perf)), dim3(TB), 0, s_, dF.as<float>(), idx, perf,
                         bF.as<float>(), bs);
      hipLaunchKernelGGL(k_gather, dim3(grid_for(long(bs) * nev_)), dim3(TB), 0, s_, dK.as<float>(), idx, long(nev_),
                         bK.as<float>(), bs);
      fwd(bs);
      bwd(bs, bF.as<float>(), bK.as<float>(), wk, e);
      ++adam_t_;
      const float c1 = float(1.0 - std::pow(double(b1), double(adam_t_)));
      const float c2 = float(1.0 - std::pow(double(b2), double(adam_t_)));
      hipLaunchKernelGGL(k_adam, dim3(grid_for(nparams_)), dim3(TB), 0, s_, P_.as<float>(), G_.as<float>(),
                         M1_.as<float>(), M2_.as<float>(), nparams_, lr, b1, b2, 1e-8f, c1, c2);
    }
  }
  OMX_HIP_CHECK(hipGetLastError());
  std::vector<double> h(epochs);
  OMX_HIP_CHECK(hipMemcpyAsync(h.data(), lossd_.as<void>(), sizeof(double) * epochs, hipMemcpyDeviceToHost, s_));
  OMX_HIP_CHECK(hipStreamSynchronize(s_));
  if (hist)
    for (int e = 0; e < epochs; ++e) hist[e] = h[e] / double(nb);
}

}  // namespace omx
