// Batched vector helpers and preconditioned MINRES (see krylov.hpp).
#include "krylov.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace omx {

namespace {

constexpr int AB = 256;

// out[(i*nb + b)] = Re sum_rows conj(a_i) b_i over block b's rows (vector i at base + i*ld),
// accumulated in FP64
template <typename T>
__global__ __launch_bounds__(AB) void k_vdot(const cplx<T>* A, const cplx<T>* B, long ld, long K, const int* act,
                                             double* part) {
  const int i = blockIdx.y;
  __shared__ double red[AB];
  double s = 0;
  if (act[i]) {
    const cplx<T>* a = A + i * ld;
    const cplx<T>* b = B + i * ld;
    for (long r = long(blockIdx.x) * AB + threadIdx.x; r < K; r += long(gridDim.x) * AB)
      s += double(a[r].x) * double(b[r].x) + double(a[r].y) * double(b[r].y);
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int t = AB / 2; t > 0; t >>= 1) {
    if (threadIdx.x < t) red[threadIdx.x] += red[threadIdx.x + t];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[long(i) * gridDim.x + blockIdx.x] = red[0];
}

// complex dot <a_i, b_i> = sum conj(a) b, partials (re, im) in FP64
template <typename T>
__global__ __launch_bounds__(AB) void k_cdot(const cplx<T>* A, const cplx<T>* B, long ld, long K, const int* act,
                                             c64* part) {
  const int i = blockIdx.y;
  __shared__ double rr[AB], ri[AB];
  double sr = 0, si = 0;
  if (act[i]) {
    const cplx<T>* a = A + i * ld;
    const cplx<T>* b = B + i * ld;
    for (long r = long(blockIdx.x) * AB + threadIdx.x; r < K; r += long(gridDim.x) * AB) {
      const double ax = a[r].x, ay = a[r].y, bx = b[r].x, by = b[r].y;
      sr += ax * bx + ay * by;
      si += ax * by - ay * bx;
    }
  }
  rr[threadIdx.x] = sr;
  ri[threadIdx.x] = si;
  __syncthreads();
  for (int t = AB / 2; t > 0; t >>= 1) {
    if (threadIdx.x < t) {
      rr[threadIdx.x] += rr[threadIdx.x + t];
      ri[threadIdx.x] += ri[threadIdx.x + t];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) part[long(i) * gridDim.x + blockIdx.x] = c64{rr[0], ri[0]};
}

// out_i = sum_j c[i*4 + j] * in_j_i for up to 4 inputs (complex FP64 coefficients, FP64
// arithmetic); out may alias an input (each element is read before it is written by the
// same thread)
template <typename T> struct Lin {
  const cplx<T>* in[4];
  int nin;
  cplx<T>* out;
  const c64* coef;  // per instance, 4 complex
  long ld;
};

template <typename T, int NIN>
__global__ __launch_bounds__(AB) void k_lincomb(Lin<T> L, long K, const int* act) {
  const int i = blockIdx.y;
  if (!act[i]) return;
  c64 cs[NIN];
  const cplx<T>* in[NIN];
#pragma unroll
  for (int j = 0; j < NIN; ++j) {
    cs[j] = L.coef[4 * i + j];
    in[j] = L.in[j] + i * L.ld;
  }
  cplx<T>* out = L.out + i * L.ld;
  for (long r = long(blockIdx.x) * AB + threadIdx.x; r < K; r += long(gridDim.x) * AB) {
    cplx<T> v[NIN];
#pragma unroll
    for (int j = 0; j < NIN; ++j) v[j] = in[j][r];
    double sx = 0, sy = 0;
#pragma unroll
    for (int j = 0; j < NIN; ++j) {
      sx += cs[j].x * double(v[j].x) - cs[j].y * double(v[j].y);
      sy += cs[j].x * double(v[j].y) + cs[j].y * double(v[j].x);
    }
    out[r] = cplx<T>{T(sx), T(sy)};
  }
}

// out_i = sum_j c_j in_j_i as k_lincomb, plus the partial sums of Re <d_i, out_i> over this
// block's rows into part[i gridDim.x + blockIdx.x] (the layout of k_vdot)
template <typename T, int NIN>
__global__ __launch_bounds__(AB) void k_lincomb_dot(Lin<T> L, long K, const int* act, const cplx<T>* D,
                                                    double* part) {
  const int i = blockIdx.y;
  __shared__ double red[AB];
  double acc = 0;
  if (act[i]) {
    c64 cs[NIN];
    const cplx<T>* in[NIN];
#pragma unroll
    for (int j = 0; j < NIN; ++j) {
      cs[j] = L.coef[4 * i + j];
      in[j] = L.in[j] + i * L.ld;
    }
    cplx<T>* out = L.out + i * L.ld;
    const cplx<T>* d = D + i * L.ld;
    for (long r = long(blockIdx.x) * AB + threadIdx.x; r < K; r += long(gridDim.x) * AB) {
      cplx<T> v[NIN];
#pragma unroll
      for (int j = 0; j < NIN; ++j) v[j] = in[j][r];
      double sx = 0, sy = 0;
#pragma unroll
      for (int j = 0; j < NIN; ++j) {
        sx += cs[j].x * double(v[j].x) - cs[j].y * double(v[j].y);
        sy += cs[j].x * double(v[j].y) + cs[j].y * double(v[j].x);
      }
      const cplx<T> o{T(sx), T(sy)};
      out[r] = o;
      acc += double(d[r].x) * double(o.x) + double(d[r].y) * double(o.y);
    }
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int t = AB / 2; t > 0; t >>= 1) {
    if (threadIdx.x < t) red[threadIdx.x] += red[threadIdx.x + t];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[long(i) * gridDim.x + blockIdx.x] = red[0];
}

// Block-wide fixed-order sum of n partials (every block of a launch gets the same value)
__device__ __forceinline__ double block_sum(const double* p, int n, double* red) {
  double a = 0;
  for (int b = threadIdx.x; b < n; b += AB) a += p[b];
  red[threadIdx.x] = a;
  __syncthreads();
  for (int t = AB / 2; t > 0; t >>= 1) {
    if (threadIdx.x < t) red[threadIdx.x] += red[threadIdx.x + t];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

struct MinresState;

// Stage-1 update fused with its dot, with v = u / beta not formed: y <- (A u - shift u) / beta
// - (beta/oldb) r1 (A u in au) and the partials of Re <v, y> = Re <u, y> / beta; the
// coefficients come from the MINRES state (read only)
//
// With WX the previous iteration's direction and solution updates come first, row by row:
// w_new = c0 v' + c1 w1' + c2 w2' (into wout) and x += c3 w_new with v' the previous
// iteration's u, which the output buffer y still holds until this row is written.
template <typename T, bool WX>
__global__ __launch_bounds__(AB) void k_minres_y1(const MinresState* ms, const cplx<T>* au, const cplx<T>* u,
                                                  const cplx<T>* r1, cplx<T>* y, long ld, long K, const int* act,
                                                  double* part, const cplx<T>* w1p, const cplx<T>* w2p,
                                                  cplx<T>* wout, cplx<T>* x, const c64* wcoef);

// Stage-2 update fused with its scalar step: alfa = sum of the <v, y> partials (every block
// the same, block 0 stores it), y <- y - (alfa/beta) r2
template <typename T>
__global__ __launch_bounds__(AB) void k_minres_y2(MinresState* ms, const double* vpart, int nvp,
                                                  const cplx<T>* r2, cplx<T>* y, long ld, long K, const int* act);

// Projection fused with its coefficient and a dot: x <- x - ev <ev, x> from the complex
// partials of <ev, x> (each block sums them), and the partials of Re <d, x_new>
template <typename T>
__global__ __launch_bounds__(AB) void k_project_dot(const c64* cpart, int ncp, const cplx<T>* ev, cplx<T>* x,
                                                    const cplx<T>* d, long ld, long K, const int* act, double* part) {
  const int i = blockIdx.y;
  __shared__ double red[AB];
  double acc = 0;
  if (act[i]) {
    // <ev, x> in fixed order: real parts, then imaginary parts (the partials interleave them)
    const double* cp = reinterpret_cast<const double*>(cpart + long(i) * ncp);
    double sr = 0, si = 0;
    for (int b = threadIdx.x; b < ncp; b += AB) sr += cp[2 * b];
    red[threadIdx.x] = sr;
    __syncthreads();
    for (int t = AB / 2; t > 0; t >>= 1) {
      if (threadIdx.x < t) red[threadIdx.x] += red[threadIdx.x + t];
      __syncthreads();
    }
    sr = red[0];
    __syncthreads();
    for (int b = threadIdx.x; b < ncp; b += AB) si += cp[2 * b + 1];
    red[threadIdx.x] = si;
    __syncthreads();
    for (int t = AB / 2; t > 0; t >>= 1) {
      if (threadIdx.x < t) red[threadIdx.x] += red[threadIdx.x + t];
      __syncthreads();
    }
    si = red[0];
    __syncthreads();
    const long o = i * ld;
    for (long r = long(blockIdx.x) * AB + threadIdx.x; r < K; r += long(gridDim.x) * AB) {
      const cplx<T> e = ev[o + r], xr = x[o + r];
      // x - ev (sr + i si)
      const double nx = double(xr.x) - (double(e.x) * sr - double(e.y) * si);
      const double ny = double(xr.y) - (double(e.x) * si + double(e.y) * sr);
      const cplx<T> xn{T(nx), T(ny)};
      x[o + r] = xn;
      acc += double(d[o + r].x) * double(xn.x) + double(d[o + r].y) * double(xn.y);
    }
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int t = AB / 2; t > 0; t >>= 1) {
    if (threadIdx.x < t) red[threadIdx.x] += red[threadIdx.x + t];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[long(i) * gridDim.x + blockIdx.x] = red[0];
}

// MINRES direction and solution updates in one pass: w_new = c0 v + c1 w1 + c2 w2 (into
// wout) and x += c3 w_new, with real coefficients c = coef[4 i ..]
template <typename T>
__global__ __launch_bounds__(AB) void k_minres_wx(const cplx<T>* v, const cplx<T>* w1, const cplx<T>* w2,
                                                  cplx<T>* wout, cplx<T>* x, long ld, long K, const c64* coef,
                                                  const int* act) {
  const int i = blockIdx.y;
  if (!act[i]) return;
  const double c0 = coef[4 * i].x, c1 = coef[4 * i + 1].x, c2 = coef[4 * i + 2].x, c3 = coef[4 * i + 3].x;
  const long o = i * ld;
  for (long r = long(blockIdx.x) * AB + threadIdx.x; r < K; r += long(gridDim.x) * AB) {
    const cplx<T> a = v[o + r], b = w1[o + r], c = w2[o + r], xr = x[o + r];
    const double wx = c0 * double(a.x) + c1 * double(b.x) + c2 * double(c.x);
    const double wy = c0 * double(a.y) + c1 * double(b.y) + c2 * double(c.y);
    wout[o + r] = cplx<T>{T(wx), T(wy)};
    x[o + r] = cplx<T>{T(double(xr.x) + c3 * wx), T(double(xr.y) + c3 * wy)};
  }
}

struct MinresState {
  double beta1, beta, oldb, dbar, epsln, phibar, cs, sn, alfa, phi, delta, oldeps, denom, rtol, shift;
  int itn, done, active, pad;
};


// Scalar recurrences of MINRES (Paige & Saunders 1975) per instance; `stage` selects the
// step of the iteration they follow.  Coefficients for the next lincomb go to coef.
// one block per instance: the block sums the instance's partials (tree), then its first
// thread runs the recurrence
template <typename T, bool WX>
__global__ __launch_bounds__(AB) void k_minres_y1(const MinresState* ms, const cplx<T>* au, const cplx<T>* u,
                                                  const cplx<T>* r1, cplx<T>* y, long ld, long K, const int* act,
                                                  double* part, const cplx<T>* w1p, const cplx<T>* w2p,
                                                  cplx<T>* wout, cplx<T>* x, const c64* wcoef) {
  const int i = blockIdx.y;
  __shared__ double red[AB];
  double acc = 0;
  if (act[i]) {
    const MinresState& m = ms[i];
    // beta = 0 ends the instance (its v is zero): the zero scale keeps its updates finite
    const double ib = m.beta > 0 ? 1.0 / m.beta : 0.0;
    const double c0 = ib, c1 = -m.shift * ib, c2 = m.itn >= 1 ? -m.beta / m.oldb : 0.0;
    double d0 = 0, d1 = 0, d2 = 0, d3 = 0;
    if (WX) {
      d0 = wcoef[4 * i].x;
      d1 = wcoef[4 * i + 1].x;
      d2 = wcoef[4 * i + 2].x;
      d3 = wcoef[4 * i + 3].x;
    }
    const long o = i * ld;
    for (long r = long(blockIdx.x) * AB + threadIdx.x; r < K; r += long(gridDim.x) * AB) {
      if (WX) {
        const cplx<T> pv = y[o + r], pb = w1p[o + r], pc = w2p[o + r], xr = x[o + r];
        const double wx = d0 * double(pv.x) + d1 * double(pb.x) + d2 * double(pc.x);
        const double wy = d0 * double(pv.y) + d1 * double(pb.y) + d2 * double(pc.y);
        wout[o + r] = cplx<T>{T(wx), T(wy)};
        x[o + r] = cplx<T>{T(double(xr.x) + d3 * wx), T(double(xr.y) + d3 * wy)};
      }
      const cplx<T> a = au[o + r], b = u[o + r], c = r1[o + r];
      const cplx<T> yn{T(c0 * double(a.x) + c1 * double(b.x) + c2 * double(c.x)),
                       T(c0 * double(a.y) + c1 * double(b.y) + c2 * double(c.y))};
      y[o + r] = yn;
      acc += double(b.x) * double(yn.x) + double(b.y) * double(yn.y);
    }
    acc *= ib;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int t = AB / 2; t > 0; t >>= 1) {
    if (threadIdx.x < t) red[threadIdx.x] += red[threadIdx.x + t];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[long(i) * gridDim.x + blockIdx.x] = red[0];
}

template <typename T>
__global__ __launch_bounds__(AB) void k_minres_y2(MinresState* ms, const double* vpart, int nvp,
                                                  const cplx<T>* r2, cplx<T>* y, long ld, long K, const int* act) {
  const int i = blockIdx.y;
  if (!act[i]) return;
  __shared__ double red[AB];
  const double alfa = block_sum(vpart + long(i) * nvp, nvp, red);
  // every block read the state it needs before block 0 writes alfa, which no block reads
  const double f = ms[i].beta > 0 ? -alfa / ms[i].beta : 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) ms[i].alfa = alfa;
  const long o = i * ld;
  for (long r = long(blockIdx.x) * AB + threadIdx.x; r < K; r += long(gridDim.x) * AB) {
    const cplx<T> a = y[o + r], c = r2[o + r];
    y[o + r] = cplx<T>{T(double(a.x) + f * double(c.x)), T(double(a.y) + f * double(c.y))};
  }
}

// Coefficient slots (4 complex per instance each, slot s at coef + 4 ninst s): 0 for the
// y updates of stages 1 and 2, 1 for the fused w/x update of stage 3, 2 for the next
// v = y / beta (stages 0 and 3); slot 3 belongs to the projections.
__global__ __launch_bounds__(AB) void k_minres_scalars(MinresState* ms, const double* part, int nb, c64* coef,
                                                       int stage, int ninst, const int* act) {
  const int i = blockIdx.x;
  __shared__ double red[AB];
  double acc = 0;
  for (int b = threadIdx.x; b < nb; b += AB) acc += part[long(i) * nb + b];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int t = AB / 2; t > 0; t >>= 1) {
    if (threadIdx.x < t) red[threadIdx.x] += red[threadIdx.x + t];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    MinresState& m = ms[i];
    c64* c = coef + 4 * i + (stage == 0 ? 8 * ninst : (stage == 3 ? 4 * ninst : 0));
    c64* cv = coef + 4 * i + 8 * ninst;  // slot 2: next v
    for (int j = 0; j < 4; ++j) c[j] = c64{0, 0};
    if (stage == 3)
      for (int j = 0; j < 4; ++j) cv[j] = c64{0, 0};
    if (!act[i]) return;
    const double d = red[0];
    if (stage == 0) {
      // init: beta1 = sqrt(<r1, M r1>)
      m.beta1 = sqrt(fmax(d, 0.0));
      m.beta = m.beta1;
      m.oldb = 0;
      m.dbar = 0;
      m.epsln = 0;
      m.phibar = m.beta1;
      m.cs = -1;
      m.sn = 0;
      m.itn = 0;
      m.done = m.beta1 == 0 ? 1 : 0;
      // next: v = y / beta
      c[0] = c64{m.done ? 0.0 : 1.0 / m.beta, 0};
    } else if (stage == 1) {
      // after y = A v: y <- y - shift v - (beta/oldb) r1 ; inputs (y, v, r1)
      c[0] = c64{1, 0};
      c[1] = c64{-m.shift, 0};
      c[2] = c64{m.itn >= 1 ? -m.beta / m.oldb : 0.0, 0};
    } else if (stage == 2) {
      // alfa = <v, y>; y <- y - (alfa/beta) r2 ; inputs (y, r2)
      m.alfa = d;
      c[0] = c64{1, 0};
      c[1] = c64{-m.alfa / m.beta, 0};
    } else if (stage == 3) {
      // beta = sqrt(<r2, M r2>); rotations; w and x updates
      m.itn += 1;
      m.oldb = m.beta;
      m.beta = sqrt(fmax(d, 0.0));
      m.oldeps = m.epsln;
      m.delta = m.cs * m.dbar + m.sn * m.alfa;
      const double gbar = m.sn * m.dbar - m.cs * m.alfa;
      m.epsln = m.sn * m.beta;
      m.dbar = -m.cs * m.beta;
      double gamma = sqrt(gbar * gbar + m.beta * m.beta);
      gamma = fmax(gamma, 2.220446049250313e-16);
      m.cs = gbar / gamma;
      m.sn = m.beta / gamma;
      m.phi = m.cs * m.phibar;
      m.phibar = m.sn * m.phibar;
      m.denom = 1.0 / gamma;
      // w_new = (v - oldeps w1 - delta w2) denom and x += phi w_new ; inputs (u, w1, w2)
      // with u = oldb v, the vector the iteration's operator ran on
      c[0] = c64{m.oldb > 0 ? m.denom / m.oldb : 0.0, 0};
      c[1] = c64{-m.oldeps * m.denom, 0};
      c[2] = c64{-m.delta * m.denom, 0};
      c[3] = c64{m.phi, 0};
      // next v = y / beta
      cv[0] = c64{m.beta > 0 ? 1.0 / m.beta : 0.0, 0};
      if (m.phibar <= m.rtol * m.beta1 || m.beta == 0) m.done = 1;
    }
  }
}

// coef[4 i] = 1, coef[4 i + 1] = -sum of the complex partials of instance i (x - ev <ev, x>);
// one block per instance
__global__ __launch_bounds__(AB) void k_proj_coef(const c64* part, int nblk, c64* coef) {
  const int i = blockIdx.x;
  __shared__ double rr[AB], ri[AB];
  double sr = 0, si = 0;
  for (int b = threadIdx.x; b < nblk; b += AB) {
    sr += part[long(i) * nblk + b].x;
    si += part[long(i) * nblk + b].y;
  }
  rr[threadIdx.x] = sr;
  ri[threadIdx.x] = si;
  __syncthreads();
  for (int t = AB / 2; t > 0; t >>= 1) {
    if (threadIdx.x < t) {
      rr[threadIdx.x] += rr[threadIdx.x + t];
      ri[threadIdx.x] += ri[threadIdx.x + t];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    coef[4 * i] = c64{1, 0};
    coef[4 * i + 1] = c64{-rr[0], -ri[0]};
    coef[4 * i + 2] = c64{0, 0};
    coef[4 * i + 3] = c64{0, 0};
  }
}

// Column loops of the block projection with the instance's column count NQ fixed at
// compile time (qdot_cols / qproject_cols pick it from the runtime count): every load of a
// row is issued before the first use, so a row costs one memory round trip, not one per column.
template <typename T, int NQ>
__device__ __forceinline__ void qdot_rows(const cplx<T>* q, long qld, const cplx<T>* x, long r0, long stride, long K,
                                          double* sr, double* si) {
  for (long r = r0; r < K; r += stride) {
    const cplx<T> xv = x[r];
    cplx<T> e[NQ];
#pragma unroll
    for (int c = 0; c < NQ; ++c) e[c] = q[c * qld + r];
    const double xx = xv.x, xy = xv.y;
#pragma unroll
    for (int c = 0; c < NQ; ++c) {
      sr[c] += double(e[c].x) * xx + double(e[c].y) * xy;
      si[c] += double(e[c].x) * xy - double(e[c].y) * xx;
    }
  }
}

template <typename T, int QM, int NQ = 1>
__device__ __forceinline__ void qdot_cols(int n, const cplx<T>* q, long qld, const cplx<T>* x, long r0, long stride,
                                          long K, double* sr, double* si) {
  if constexpr (NQ <= QM) {
    if (n == NQ)
      qdot_rows<T, NQ>(q, qld, x, r0, stride, K, sr, si);
    else
      qdot_cols<T, QM, NQ + 1>(n, q, qld, x, r0, stride, K, sr, si);
  }
}

// x <- x + sum_c cf_c Q_c over the rows, and sum of Re <d, x_new> when DOT
template <typename T, int NQ, bool DOT>
__device__ __forceinline__ double qproject_rows(const c64* cf, const cplx<T>* q, long qld, cplx<T>* x,
                                                const cplx<T>* dv, long r0, long stride, long K) {
  double acc = 0;
  for (long r = r0; r < K; r += stride) {
    const cplx<T> xv = x[r];
    cplx<T> dr{T(0), T(0)};
    if (DOT) dr = dv[r];
    cplx<T> e[NQ];
#pragma unroll
    for (int c = 0; c < NQ; ++c) e[c] = q[c * qld + r];
    double nx = xv.x, ny = xv.y;
#pragma unroll
    for (int c = 0; c < NQ; ++c) {
      nx += cf[c].x * double(e[c].x) - cf[c].y * double(e[c].y);
      ny += cf[c].x * double(e[c].y) + cf[c].y * double(e[c].x);
    }
    const cplx<T> xn{T(nx), T(ny)};
    x[r] = xn;
    if (DOT) acc += double(dr.x) * double(xn.x) + double(dr.y) * double(xn.y);
  }
  return acc;
}

template <typename T, int QM, int NQ = 1>
__device__ __forceinline__ double qproject_cols(int n, const c64* cf, const cplx<T>* q, long qld, cplx<T>* x,
                                                const cplx<T>* dv, long r0, long stride, long K) {
  if constexpr (NQ <= QM) {
    if (n == NQ)
      return dv ? qproject_rows<T, NQ, true>(cf, q, qld, x, dv, r0, stride, K)
                : qproject_rows<T, NQ, false>(cf, q, qld, x, dv, r0, stride, K);
    return qproject_cols<T, QM, NQ + 1>(n, cf, q, qld, x, dv, r0, stride, K);
  }
  return 0.0;
}

// Block projection: partials of <Q_c, x> (c < nq[i], complex, FP64) over this block's
// rows into part[(i gridDim.x + blockIdx.x) QM + c] (zero for c >= nq[i])
template <typename T, int QM>
__global__ __launch_bounds__(AB) void k_qdot(const cplx<T>* Q, long qld, long qis, const int* nq, const cplx<T>* X,
                                             long ld, long K, const int* act, c64* part) {
  const int i = blockIdx.y;
  const int n = act[i] ? nq[i] : 0;
  __shared__ double wr[AB / 64][QM], wi[AB / 64][QM];
  double sr[QM], si[QM];
#pragma unroll
  for (int c = 0; c < QM; ++c) sr[c] = si[c] = 0.0;
  if (n > 0)
    qdot_cols<T, QM>(n, Q + i * qis, qld, X + i * ld, long(blockIdx.x) * AB + threadIdx.x, long(gridDim.x) * AB, K,
                     sr, si);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < QM; ++c) {
    double a = sr[c], b = si[c];
    for (int off = 32; off > 0; off >>= 1) {
      a += __shfl_down(a, off, 64);
      b += __shfl_down(b, off, 64);
    }
    if (lane == 0) {
      wr[wv][c] = a;
      wi[wv][c] = b;
    }
  }
  __syncthreads();
  if (threadIdx.x < QM) {
    const int c = threadIdx.x;
    double a = 0, b = 0;
    for (int w = 0; w < AB / 64; ++w) {
      a += wr[w][c];
      b += wi[w][c];
    }
    part[(long(i) * gridDim.x + blockIdx.x) * QM + c] = c64{a, b};
  }
}

// x_i <- x_i - sum_c <Q_c, x_i> Q_c (the block projection), the coefficients summed by every
// block from k_qdot's qnblk partials of its instance in one fixed order; with d, the
// partials of Re <d_i, x_i new> into part[i gridDim.x + blockIdx.x]
template <typename T, int QM>
__global__ __launch_bounds__(AB) void k_qproject(const c64* qpart, int qnblk, const cplx<T>* Q, long qld, long qis,
                                                 const int* nq, cplx<T>* X, const cplx<T>* d, long ld, long K,
                                                 const int* act, double* part) {
  constexpr int G = AB / QM;
  const int i = blockIdx.y;
  __shared__ double red[AB];
  __shared__ double gr[G][QM], gi[G][QM];
  __shared__ c64 cfs[QM];
  double acc = 0;
  if (act[i]) {
    const int n = nq[i];
    {
      const int c = threadIdx.x % QM, j = threadIdx.x / QM;
      double a = 0, b = 0;
      for (int q = j; q < qnblk; q += G) {
        const c64 v = qpart[(long(i) * qnblk + q) * QM + c];
        a += v.x;
        b += v.y;
      }
      gr[j][c] = a;
      gi[j][c] = b;
      __syncthreads();
      if (threadIdx.x < QM) {
        double sa = 0, sb = 0;
        for (int g = 0; g < G; ++g) {
          sa += gr[g][threadIdx.x];
          sb += gi[g][threadIdx.x];
        }
        cfs[threadIdx.x] = c64{-sa, -sb};
      }
      __syncthreads();
    }
    c64 cf[QM];
#pragma unroll
    for (int c = 0; c < QM; ++c) cf[c] = c < n ? cfs[c] : c64{0, 0};
    if (n > 0)
      acc = qproject_cols<T, QM>(n, cf, Q + i * qis, qld, X + i * ld, d ? d + i * ld : nullptr,
                                 long(blockIdx.x) * AB + threadIdx.x, long(gridDim.x) * AB, K);
  }
  if (!part) return;
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int t = AB / 2; t > 0; t >>= 1) {
    if (threadIdx.x < t) red[threadIdx.x] += red[threadIdx.x + t];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[long(i) * gridDim.x + blockIdx.x] = red[0];
}

}  // namespace

double sum_part(const double* p, int nb) {
  double s = 0;
  for (int b = 0; b < nb; ++b) s += p[b];
  return s;
}

template <typename T>
VecOpsT<T>::VecOpsT(Solver& S, hipStream_t s_, int nb_, long ld_, long K_, const C* EV_)
    : s(s_), nb(nb_), ld(ld_), K(K_), nblk(int(std::min<long>(256, (K_ + AB - 1) / AB))), EV(EV_),
      part(S.vec_part), coef(S.vec_coef) {
  part.ensure(sizeof(c64) * size_t(nb_) * 256);
  coef.ensure(sizeof(c64) * 4 * 4 * size_t(nb_));  // four coefficient slots
}

template <typename T>
void VecOpsT<T>::lincomb_dot(std::initializer_list<const C*> ins, C* outp, const int* act, const C* d, int slot) {
  Lin<T> L{};
  int j = 0;
  for (auto p : ins) L.in[j++] = p;
  L.nin = j;
  L.out = outp;
  L.coef = coef.template as<c64>() + 4 * size_t(nb) * slot;
  L.ld = ld;
  const dim3 g(nblk, nb);
  switch (L.nin) {
    case 2: hipLaunchKernelGGL((k_lincomb_dot<T, 2>), g, dim3(AB), 0, s, L, K, act, d, part.template as<double>()); break;
    case 3: hipLaunchKernelGGL((k_lincomb_dot<T, 3>), g, dim3(AB), 0, s, L, K, act, d, part.template as<double>()); break;
    default: throw Error(OMX_ERR_INTERNAL, "lincomb_dot takes 2 or 3 inputs");
  }
}

template <typename T> void VecOpsT<T>::vdot(const C* A, const C* B, const int* act) {
  hipLaunchKernelGGL(k_vdot<T>, dim3(nblk, nb), dim3(AB), 0, s, A, B, ld, K, act, part.as<double>());
}

template <typename T>
void VecOpsT<T>::lincomb(std::initializer_list<const C*> ins, C* outp, const int* act, int slot) {
  Lin<T> L{};
  int j = 0;
  for (auto p : ins) L.in[j++] = p;
  L.nin = j;
  L.out = outp;
  L.coef = coef.template as<c64>() + 4 * size_t(nb) * slot;
  L.ld = ld;
  const dim3 g(nblk, nb);
  switch (L.nin) {
    case 1: hipLaunchKernelGGL((k_lincomb<T, 1>), g, dim3(AB), 0, s, L, K, act); break;
    case 2: hipLaunchKernelGGL((k_lincomb<T, 2>), g, dim3(AB), 0, s, L, K, act); break;
    case 3: hipLaunchKernelGGL((k_lincomb<T, 3>), g, dim3(AB), 0, s, L, K, act); break;
    default: hipLaunchKernelGGL((k_lincomb<T, 4>), g, dim3(AB), 0, s, L, K, act); break;
  }
}

template <typename T> void VecOpsT<T>::set_block(const C* Qb, long qis, const int* nq, int nq_max) {
  if (nq_max < 1 || nq_max > QMAX) throw Error(OMX_ERR_INTERNAL, "block projection: 1..16 columns");
  // the real dot partials (nb nblk, at the start, where the MINRES steps read them), then
  // the block partials (nb nblk QMAX complex): every block of k_qproject reads the latter
  // while others write the former, so the two never overlap
  part.ensure(sizeof(c64) * size_t(nb) * (size_t(nblk) + size_t(nblk) * QMAX));
  Q = Qb;
  qistride = qis;
  qn = nq;
  qmax = nq_max;
  // a few thousand rows per block: the per-block reduction of the column dots stays small
  // beside the rows' loads
  qblk = std::max(1, std::min(nblk, 64));
}

template <typename T> void VecOpsT<T>::project_dot_to(C* x, const int* act, const C* dot_with, double* out) {
  const dim3 g(nblk, nb);
  if (Q) {
    c64* qp = part.template as<c64>() + size_t(nb) * nblk;
    const dim3 gq(qblk, nb);
    if (qmax <= 8) {
      hipLaunchKernelGGL((k_qdot<T, 8>), gq, dim3(AB), 0, s, Q, ld, qistride, qn, x, ld, K, act, qp);
      hipLaunchKernelGGL((k_qproject<T, 8>), g, dim3(AB), 0, s, qp, qblk, Q, ld, qistride, qn, x, dot_with, ld, K,
                         act, out);
    } else {
      hipLaunchKernelGGL((k_qdot<T, QMAX>), gq, dim3(AB), 0, s, Q, ld, qistride, qn, x, ld, K, act, qp);
      hipLaunchKernelGGL((k_qproject<T, QMAX>), g, dim3(AB), 0, s, qp, qblk, Q, ld, qistride, qn, x, dot_with, ld, K,
                         act, out);
    }
    return;
  }
  hipLaunchKernelGGL(k_cdot<T>, g, dim3(AB), 0, s, EV, x, ld, K, act, part.template as<c64>());
  hipLaunchKernelGGL((k_project_dot<T>), g, dim3(AB), 0, s, part.template as<c64>(), nblk, EV, x, dot_with, ld, K,
                     act, out);
}

template <typename T> void VecOpsT<T>::project(C* x, const int* act, const C* dot_with) {
  if (Q) {
    // the real partials land at the start of part, past which the block coefficients sit
    project_dot_to(x, act, dot_with, dot_with ? part.template as<double>() : nullptr);
    return;
  }
  hipLaunchKernelGGL(k_cdot<T>, dim3(nblk, nb), dim3(AB), 0, s, EV, x, ld, K, act, part.template as<c64>());
  hipLaunchKernelGGL(k_proj_coef, dim3(nb), dim3(AB), 0, s, part.template as<c64>(), nblk,
                     coef.template as<c64>() + 4 * size_t(nb) * 3);
  if (dot_with)
    lincomb_dot({x, EV}, x, act, dot_with, 3);
  else
    lincomb({x, EV}, x, act, 3);
}

template <typename T>
int minres_complement_t(Solver& S, VecOpsT<T>& V, const std::vector<double>& shifts, const double* dk, cplx<T>* r,
                        cplx<T>* xout, const std::vector<int>& want, double rtol, int maxit, bool throw_on_cap) {
  using C = cplx<T>;
  hipStream_t s = S.stream;
  const int nb = V.nb;
  const long ld = S.ld, n3 = 3 * S.N;
  // workspace kept on the handle (grown on demand): 8 vectors and the operator scratch per
  // instance, the MINRES states and two flag arrays
  const size_t vb = sizeof(C) * nb * ld;
  const int iblk = ((nb + 63) / 64) * 64;  // ints per flag array
  const size_t need = 8 * vb + sizeof(C) * nb * n3 + sizeof(MinresState) * nb + 2 * sizeof(int) * iblk +
                      sizeof(double) * nb * V.nblk;
  S.krylov_ws.ensure(need);
  char* wp = S.krylov_ws.as<char>();
  C *x = reinterpret_cast<C*>(wp), *r1 = r, *r2 = reinterpret_cast<C*>(wp + vb),
    *y = reinterpret_cast<C*>(wp + 2 * vb), *v = reinterpret_cast<C*>(wp + 3 * vb),
    *w = reinterpret_cast<C*>(wp + 4 * vb), *w1 = reinterpret_cast<C*>(wp + 5 * vb),
    *w2 = reinterpret_cast<C*>(wp + 6 * vb), *avp = reinterpret_cast<C*>(wp + 7 * vb);
  C* dm = reinterpret_cast<C*>(wp + 8 * vb);
  MinresState* ms = reinterpret_cast<MinresState*>(wp + 8 * vb + sizeof(C) * nb * n3);
  int* actp = reinterpret_cast<int*>(reinterpret_cast<char*>(ms) + sizeof(MinresState) * nb);
  int* livep = actp + iblk;
  // real dot partials of the fused projection (its complex input partials live in V.part)
  double* cpart2 = reinterpret_cast<double*>(livep + iblk);
  OMX_HIP_CHECK(hipMemcpyAsync(actp, want.data(), sizeof(int) * nb, hipMemcpyHostToDevice, s));
  const int* act = actp;
  OMX_HIP_CHECK(hipMemsetAsync(x, 0, vb, s));
  OMX_HIP_CHECK(hipMemsetAsync(w, 0, vb, s));
  OMX_HIP_CHECK(hipMemsetAsync(w2, 0, vb, s));
  OMX_HIP_CHECK(hipMemsetAsync(w1, 0, vb, s));
  std::vector<MinresState> hm(nb);
  for (int i = 0; i < nb; ++i) {
    std::memset(&hm[i], 0, sizeof(MinresState));
    hm[i].rtol = rtol;
    hm[i].shift = shifts[i];
    hm[i].active = want[i];
  }
  OMX_HIP_CHECK(hipMemcpyAsync(ms, hm.data(), sizeof(MinresState) * nb, hipMemcpyHostToDevice, s));
  const int nblk = V.nblk;
  double* part = V.part.template as<double>();
  c64* coef = V.coef.template as<c64>();
  V.project(r1, act);
  BView<T> AVv{avp, ld, ld}, Yv{y, ld, ld}, R1v{r1, ld, ld};
  const VecSel sel{nb, 1, nullptr, nullptr};
  const OpPlan<T>& plan = S.pd<T>().plan;
  // y = P r1 (projected)
  launch_operator<T>(OP_P, plan, R1v, Yv, dm, sel, dk, nullptr, 0, s);
  V.project(y, act, r1);  // and the partials of <r1, y>
  hipLaunchKernelGGL(k_minres_scalars, dim3(nb), dim3(AB), 0, s, ms, part, nblk, coef, 0, nb, act);
  OMX_HIP_CHECK(hipMemcpyAsync(r2, r1, vb, hipMemcpyDeviceToDevice, s));
  std::vector<int> live(want);
  int it = 0;
  const int* la = livep;
  const dim3 vg(nblk, nb);
  // One MINRES iteration on u = y (beta v; v itself is never formed).  Its buffer rotations
  // are a 4-cycle (y <- r1 <- r2 <- v <- y) and a 3-cycle (w <- w1 <- w2 <- w), so twelve
  // iterations leave every pointer where it was.
  // The direction and solution updates of an iteration (w_new = (v - oldeps w1' - delta w2')
  // denom with w1' = w2 and w2' = w, MINRES's w1 <- w2, w2 <- w, written into the buffer of
  // the discarded w1, and x += phi w_new, with v = u / oldb folded into the coefficient)
  // wait, pending, for the next iteration's first update, which runs them row by row before
  // its own (k_minres_y1<T, true>), or for flush at the end of a group.
  C *pw1 = nullptr, *pw2 = nullptr, *pwo = nullptr;
  auto iteration = [&](bool fuse) {
    BView<T> Uv{y, ld, ld};
    launch_operator<T>(OP_M, plan, Uv, AVv, dm, sel, dk, nullptr, 0, s);
    // v-buffer <- (A u - shift u) / beta - (beta/oldb) r1 with the partials of <v, .>
    if (fuse)
      hipLaunchKernelGGL((k_minres_y1<T, true>), vg, dim3(AB), 0, s, ms, avp, y, r1, v, ld, V.K, la, part, pw1, pw2,
                         pwo, x, coef + 4 * nb);
    else
      hipLaunchKernelGGL((k_minres_y1<T, false>), vg, dim3(AB), 0, s, ms, avp, y, r1, v, ld, V.K, la, part, nullptr,
                         nullptr, nullptr, nullptr, nullptr);
    // alfa = <v, .>; v-buffer -= (alfa/beta) r2 (stage 2 fused)
    hipLaunchKernelGGL((k_minres_y2<T>), vg, dim3(AB), 0, s, ms, part, nblk, r2, v, ld, V.K, la);
    C* u = y;   // the next iteration's v-buffer, read first by this iteration's w/x update
    C* yn = r1;  // free: receives P r2 of the new r2
    r1 = r2;
    r2 = v;
    BView<T> R2b{r2, ld, ld}, Yb{yn, ld, ld};
    launch_operator<T>(OP_P, plan, R2b, Yb, dm, sel, dk, nullptr, 0, s);
    // yn <- yn - ev <ev, yn> (or the block projection) with the partials of <r2, yn>
    V.project_dot_to(yn, la, r2, cpart2);
    hipLaunchKernelGGL(k_minres_scalars, dim3(nb), dim3(AB), 0, s, ms, cpart2, nblk, coef, 3, nb, la);
    // this iteration's w/x update, pending (its v is u, the next iteration's v-buffer)
    pw1 = w2;
    pw2 = w;
    pwo = w1;
    {
      C* wn = w1;
      w1 = w2;
      w2 = w;
      w = wn;
    }
    v = u;
    y = yn;
  };
  auto flush = [&]() {
    hipLaunchKernelGGL((k_minres_wx<T>), vg, dim3(AB), 0, s, v, pw1, pw2, pwo, x, ld, V.K, coef + 4 * nb, la);
  };
  // The host reads the MINRES states every three iterations; with option "graphs" groups
  // of twelve (the rotations' common period) are replayed from one captured HIP graph
  // instead: the kernels are short and the host's launch rate, not the GPU, sets their
  // pace.  The graph holds kernels only (no memset or copy nodes).
  const int GROUP = S.opt.graphs ? 12 : 3;
  hipGraphExec_t gexec = nullptr;
  for (; it < maxit; it += GROUP) {
    OMX_HIP_CHECK(hipMemcpyAsync(hm.data(), ms, sizeof(MinresState) * nb, hipMemcpyDeviceToHost, s));
    OMX_HIP_CHECK(hipStreamSynchronize(s));
    bool anyl = false;
    for (int i = 0; i < nb; ++i) {
      live[i] = want[i] && !hm[i].done;
      anyl = anyl || live[i];
    }
    if (!anyl) break;
    OMX_HIP_CHECK(hipMemcpyAsync(livep, live.data(), sizeof(int) * nb, hipMemcpyHostToDevice, s));
    const int nit = std::min(GROUP, maxit - it);
    if (S.opt.graphs && nit == GROUP) {
      if (!gexec) {
        hipGraph_t g = nullptr;
        OMX_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        for (int u = 0; u < GROUP; ++u) iteration(u > 0);
        flush();
        OMX_HIP_CHECK(hipStreamEndCapture(s, &g));
        const hipError_t e = hipGraphInstantiate(&gexec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        OMX_HIP_CHECK(e);
      }
      OMX_HIP_CHECK(hipGraphLaunch(gexec, s));
    } else {
      for (int u = 0; u < nit; ++u) iteration(u > 0);
      flush();
    }
  }
  if (gexec) (void)hipGraphExecDestroy(gexec);
  OMX_HIP_CHECK(hipMemcpyAsync(hm.data(), ms, sizeof(MinresState) * nb, hipMemcpyDeviceToHost, s));
  OMX_HIP_CHECK(hipStreamSynchronize(s));
  if (throw_on_cap)
    for (int i = 0; i < nb; ++i)
      if (want[i] && !hm[i].done) throw Error(OMX_ERR_NOT_CONVERGED, "MINRES did not converge within the iteration cap");
  OMX_HIP_CHECK(hipMemcpyAsync(xout, x, vb, hipMemcpyDeviceToDevice, s));
  V.project(xout, act);
  V.project(xout, act);
  return it;
}

template struct VecOpsT<double>;
template struct VecOpsT<float>;
template int minres_complement_t<double>(Solver&, VecOpsT<double>&, const std::vector<double>&, const double*, c64*,
                                         c64*, const std::vector<int>&, double, int, bool);
template int minres_complement_t<float>(Solver&, VecOpsT<float>&, const std::vector<double>&, const double*,
                                        cplx<float>*, cplx<float>*, const std::vector<int>&, double, int, bool);


}  // namespace omx
