// Pull-back of rrule(solve_k) on the GPU, FP64 throughout (lib/MaxwellEigenmodes/src/
// solve.jl:213-289; device twin solvers/gpu.jl:292-365).  Per band i with primal (k_i, ev_i)
// and cotangents (kbar_i, evbar_i):
//   slope = 2 HMkH(ev)
//   if evbar != 0:
//     lambda: (M - w^2) lambda = evbar - ev <ev, evbar>, lambda ⟂ ev  (eig_adjt,
//             grads/solve.jl:96-113) -- here preconditioned MINRES on the complement of ev
//             (M - w^2 is Hermitian, singular along ev and indefinite above band 1)
//     d = fft(kx_tc(ev))/N, ld = fft(kx_tc(lambda)); epsbar += epsinv_bar(d, ld)
//     le = bfft(eps^-1 ld / N), e = bfft(eps^-1 d); khbar = -grad_k(mag, m, n)(...)
//   s = (kbar + khbar) / slope; epsbar += epsinv_bar(d, fft(kx_tc(s ev))); wbar += 2 w s
#include <algorithm>
#include <cmath>
#include <cstring>

#include "krylov.hpp"

namespace omx {

namespace {

constexpr int AB = 256;

// epsbar (6 per pixel: xx, yy, zz, xy, xz, yz) += sum over instances of epsinv_bar(d, c_i * ld)
// with the 2D mirrored convention (grads/solve.jl:490-515).  self_mode: ld = d.
__global__ __launch_bounds__(AB) void k_epsbar(const c64* Dd, const c64* Dl, long n3, long N, int ninst, const int* act,
                                               const double* cfac, int self_mode, double* eb6) {
  for (long p = long(blockIdx.x) * AB + threadIdx.x; p < N; p += long(gridDim.x) * AB) {
    double a[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < ninst; ++i) {
      if (!act[i]) continue;
      const c64* d = Dd + i * n3 + 3 * p;
      const c64* l = (self_mode ? Dd : Dl) + i * n3 + 3 * p;
      const double c = cfac ? cfac[i] : 1.0;
      // -Re(l_a conj d_a)
      for (int q = 0; q < 3; ++q) a[q] -= c * (l[q].x * d[q].x + l[q].y * d[q].y);
      // (i,j): -Re(l_j conj d_i + l_i conj d_j)
      const int ii[3] = {0, 0, 1}, jj[3] = {1, 2, 2};
      for (int q = 0; q < 3; ++q) {
        const c64 li = l[ii[q]], lj = l[jj[q]], di = d[ii[q]], dj = d[jj[q]];
        a[3 + q] -= c * ((lj.x * di.x + lj.y * di.y) + (li.x * dj.x + li.y * dj.y));
      }
    }
    for (int q = 0; q < 6; ++q) eb6[6 * p + q] += a[q];
  }
}

// k̄ from the basis sensitivities (maxwell.jl:560-566 ∇ₖmag_m_n with dk = z, as used at
// solve.jl:258-270): per G, kx̄_m = Re(le conj(ev2) + e conj(l2)), kx̄_n = -Re(le conj(ev1)
// + e conj(l1)), māg = n.kx̄_n + m.kx̄_m; partial sums of -(k̄_mag + k̄_m + k̄_n), i.e. of
// -w̄_z, the z-component of the plane wave's cotangent of w = k_vec - G (grads/period.jl:
// 97-119).  On 3D cells (nz > 1) plane iz has w_z = k - g_z and partL receives the partial
// sums of -w̄_z g_z / Dz (Λ̄_H, period.jl:209-210).
__global__ __launch_bounds__(AB) void k_khbar(const c64* LE, const c64* EE, const c64* EV, const c64* LA, long n3,
                                              long ld, int nx, int ny, int nz, double invdx, double invdy,
                                              double invdz, const double* kinst, const int* act, double* part,
                                              double* partL) {
  const int i = blockIdx.y;
  __shared__ double red[AB], redL[AB];
  double acc = 0, accL = 0;
  if (act[i]) {
    const double k = kinst[i];
    const long nxy = long(nx) * ny, N = nxy * nz;
    for (long p = long(blockIdx.x) * AB + threadIdx.x; p < N; p += long(gridDim.x) * AB) {
      const int iz = int(p / nxy);
      const long pxy = p - long(iz) * nxy;
      const int ix = int(pxy % nx), iy = int(pxy / nx);
      const double gx = double(fft_index(ix, nx)) * invdx, gy = double(fft_index(iy, ny)) * invdy;
      const double gz = nz > 1 ? double(fft_index(iz, nz)) * invdz : 0.0;
      const double wx = -gx, wy = -gy, wz = k - gz;
      const double mag = sqrt(wx * wx + wy * wy + wz * wz);
      double n0, n1;
      const double rho2 = wx * wx + wy * wy;
      if (rho2 > 0) {
        const double ir = 1.0 / sqrt(rho2);
        n0 = -wy * ir;
        n1 = wx * ir;
      } else {
        n0 = 0;
        n1 = 1;
      }
      const double n2 = 0;
      double m0 = n1 * wz - n2 * wy, m1 = n2 * wx - n0 * wz, m2 = n0 * wy - n1 * wx;
      const double im = 1.0 / sqrt(m0 * m0 + m1 * m1 + m2 * m2);
      m0 *= im;
      m1 *= im;
      m2 *= im;
      const c64* le = LE + i * n3 + 3 * p;
      const c64* ee = EE + i * n3 + 3 * p;
      const c64 e1 = EV[i * ld + 2 * p], e2 = EV[i * ld + 2 * p + 1];
      const c64 l1 = LA[i * ld + 2 * p], l2 = LA[i * ld + 2 * p + 1];
      double km[3], kn[3];
      for (int a = 0; a < 3; ++a) {
        // Re(x conj(y)) = x.re y.re + x.im y.im
        km[a] = (le[a].x * e2.x + le[a].y * e2.y) + (ee[a].x * l2.x + ee[a].y * l2.y);
        kn[a] = -((le[a].x * e1.x + le[a].y * e1.y) + (ee[a].x * l1.x + ee[a].y * l1.y));
      }
      const double magbar = n0 * kn[0] + n1 * kn[1] + n2 * kn[2] + m0 * km[0] + m1 * km[1] + m2 * km[2];
      // kpg_over_mag = (m x n) / mag
      const double c0 = (m1 * n2 - m2 * n1) / mag, c1 = (m2 * n0 - m0 * n2) / mag, c2 = (m0 * n1 - m1 * n0) / mag;
      double t = magbar * c2 * mag;
      // q = kpg_over_mag x z = (c1, -c0, 0)
      const double q0 = c1, q1 = -c0, q2 = 0;
      const double mq0 = m1 * q2 - m2 * q1, mq1 = m2 * q0 - m0 * q2, mq2 = m0 * q1 - m1 * q0;
      const double nq0 = n1 * q2 - n2 * q1, nq1 = n2 * q0 - n0 * q2, nq2 = n0 * q1 - n1 * q0;
      t -= mag * (km[0] * mq0 + km[1] * mq1 + km[2] * mq2);
      t -= mag * (kn[0] * nq0 + kn[1] * nq1 + kn[2] * nq2);
      acc -= t;
      accL -= t * gz * invdz;
    }
  }
  red[threadIdx.x] = acc;
  redL[threadIdx.x] = accL;
  __syncthreads();
  for (int s = AB / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      red[threadIdx.x] += red[threadIdx.x + s];
      redL[threadIdx.x] += redL[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[long(i) * gridDim.x + blockIdx.x] = red[0];
    if (partL) partL[long(i) * gridDim.x + blockIdx.x] = redL[0];
  }
}

// out[i*nb + b] partials of Re sum_p X_p^H A_p Y_p over the pixels (3-component fields in
// the [v][iy][ix][c] layout, A symmetric as 6 entries xx yy zz xy xz yz per pixel)
__global__ __launch_bounds__(AB) void k_qform(const c64* X, const c64* Y, const double* A6, long n3, long N,
                                              double* part) {
  const int i = blockIdx.y;
  __shared__ double red[AB];
  double acc = 0;
  const c64* x = X + i * n3;
  const c64* y = Y + i * n3;
  for (long p = long(blockIdx.x) * AB + threadIdx.x; p < N; p += long(gridDim.x) * AB) {
    const double* a = A6 + 6 * p;
    const c64 y0 = y[3 * p], y1 = y[3 * p + 1], y2 = y[3 * p + 2];
    const c64 x0 = x[3 * p], x1 = x[3 * p + 1], x2 = x[3 * p + 2];
    const double z0r = a[0] * y0.x + a[3] * y1.x + a[4] * y2.x, z0i = a[0] * y0.y + a[3] * y1.y + a[4] * y2.y;
    const double z1r = a[3] * y0.x + a[1] * y1.x + a[5] * y2.x, z1i = a[3] * y0.y + a[1] * y1.y + a[5] * y2.y;
    const double z2r = a[4] * y0.x + a[5] * y1.x + a[2] * y2.x, z2i = a[4] * y0.y + a[5] * y1.y + a[2] * y2.y;
    acc += (x0.x * z0r + x0.y * z0i) + (x1.x * z1r + x1.y * z1i) + (x2.x * z2r + x2.y * z2i);
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int t = AB / 2; t > 0; t >>= 1) {
    if (threadIdx.x < t) red[threadIdx.x] += red[threadIdx.x + t];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[long(i) * gridDim.x + blockIdx.x] = red[0];
}

// out_i = c0 kx_ct(F1_i) + c1 zx_ct(F2_i) + c2 Y_i + c3 H_i with per-instance coefficients
// c = cf[4 i ..]; F1, F2 are 3-component k-space fields ([v][iy][ix][c]), kx_ct(e) =
// (-(e.n) mag, (e.m) mag) and zx_ct(e) = (-e_x m_y + e_y m_x, -e_x n_y + e_y n_x)
// (maxwell.jl:209-216, :231-240).
__global__ __launch_bounds__(AB) void k_ct_combine(const c64* F1, const c64* F2, const c64* Y, const c64* H,
                                                   long n3, long ld, int nx, int ny, double invdx, double invdy,
                                                   const double* kinst, const double* cf, c64* out) {
  const int i = blockIdx.y;
  const double c0 = cf[4 * i], c1 = cf[4 * i + 1], c2 = cf[4 * i + 2], c3 = cf[4 * i + 3];
  const double k = kinst[i];
  const long N = long(nx) * ny;
  for (long p = long(blockIdx.x) * AB + threadIdx.x; p < N; p += long(gridDim.x) * AB) {
    const int ix = int(p % nx), iy = int(p / nx);
    const Basis<double> b = basis_at<double>(ix, iy, nx, ny, invdx, invdy, k);
    const c64* e = F1 + i * n3 + 3 * p;
    const c64* f = F2 + i * n3 + 3 * p;
    const c64 en{e[0].x * b.n0 + e[1].x * b.n1, e[0].y * b.n0 + e[1].y * b.n1};
    const c64 em{e[0].x * b.m0 + e[1].x * b.m1 + e[2].x * b.m2, e[0].y * b.m0 + e[1].y * b.m1 + e[2].y * b.m2};
    const c64 z1{-f[0].x * b.m1 + f[1].x * b.m0, -f[0].y * b.m1 + f[1].y * b.m0};
    const c64 z2{-f[0].x * b.n1 + f[1].x * b.n0, -f[0].y * b.n1 + f[1].y * b.n0};
    const long o = i * ld + 2 * p;
    const c64 y1 = Y[o], y2 = Y[o + 1], h1 = H[o], h2 = H[o + 1];
    out[o] = c64{-c0 * b.mag * en.x + c1 * z1.x + c2 * y1.x + c3 * h1.x,
                 -c0 * b.mag * en.y + c1 * z1.y + c2 * y1.y + c3 * h1.y};
    out[o + 1] = c64{c0 * b.mag * em.x + c1 * z2.x + c2 * y2.x + c3 * h2.x,
                     c0 * b.mag * em.y + c1 * z2.y + c2 * y2.y + c3 * h2.y};
  }
}

}  // namespace

void Solver::adjoint(double omega, int nev, const double* kmags, const double* evecs, const double* kbar,
                     const double* evbar, const uint8_t* evbar_mask, double* omegabar, double* epsinvbar,
                     double* Lbar) {
  if (Lbar && nz <= 1) throw Error(OMX_ERR_UNSUPPORTED, "the period cotangent needs a 3D cell (Nz > 1)");
  if (nev < 1 || nev > 64) throw Error(OMX_ERR_UNSUPPORTED, "nev must be in 1..64");
  if (!kmags || !evecs || !omegabar || !epsinvbar) throw Error(OMX_ERR_INVALID_ARG, "NULL argument");
  if (!(omega > 0) || !std::isfinite(omega)) throw Error(OMX_ERR_INVALID_ARG, "omega must be positive and finite");
  for (int i = 0; i < nev; ++i)
    if (!std::isfinite(kmags[i])) throw Error(OMX_ERR_INVALID_ARG, "kmags must be finite");
  OMX_HIP_CHECK(hipSetDevice(device));
  hipStream_t s = stream;
  const int nb = nev;
  const long n3 = 3 * N;
  const int nblk = int(std::min<long>(256, (K + AB - 1) / AB));
  DevMem dk(sizeof(double) * nb), EV(sizeof(c64) * nb * ld), EB(sizeof(c64) * nb * ld), LA(sizeof(c64) * nb * ld),
      Dd(sizeof(c64) * nb * n3), Dl(sizeof(c64) * nb * n3), eb6(sizeof(double) * 6 * N), actd(sizeof(int) * nb),
      part(sizeof(c64) * nb * nblk), partL(sizeof(double) * nb * nblk), cfac(sizeof(double) * nb);
  OMX_HIP_CHECK(hipMemsetAsync(eb6.as<void>(), 0, sizeof(double) * 6 * N, s));
  OMX_HIP_CHECK(hipMemcpyAsync(dk.as<void>(), kmags, sizeof(double) * nb, hipMemcpyHostToDevice, s));
  for (int i = 0; i < nb; ++i)
    OMX_HIP_CHECK(hipMemcpyAsync(EV.as<c64>() + i * ld, evecs + 2 * size_t(i) * K, sizeof(c64) * K,
                                 hipMemcpyHostToDevice, s));
  std::vector<int> all(nb, 1), hasv(nb, 0);
  bool any_ev = false;
  if (evbar) {
    for (int i = 0; i < nb; ++i) {
      hasv[i] = evbar_mask ? (evbar_mask[i] != 0) : 1;
      any_ev = any_ev || hasv[i];
    }
  }
  const int dps = p64.plan.dot_blocks;
  // slope = 2 HMkH(ev) for every band
  DevMem hdots(sizeof(double) * nb * dps);
  VecSel sel1{nb, 1, nullptr, nullptr};
  BView<double> EVv{EV.as<c64>(), ld, ld};
  launch_operator<double>(OP_MKH, p64.plan, EVv, EVv, Dd.as<c64>(), sel1, dk.as<double>(), hdots.as<double>(), dps, s);
  std::vector<double> hh(size_t(nb) * dps);
  OMX_HIP_CHECK(hipMemcpyAsync(hh.data(), hdots.as<void>(), sizeof(double) * hh.size(), hipMemcpyDeviceToHost, s));
  OMX_HIP_CHECK(hipStreamSynchronize(s));
  std::vector<double> slope(nb), khb(nb, 0.0), dL(nb, 0.0), LHb(nb, 0.0);
  for (int i = 0; i < nb; ++i) slope[i] = 2.0 * sum_part(hh.data() + size_t(i) * dps, dps);
  if (Lbar) {
    // dω²/dΛ = 2 HMkH_weighted(ev, g_z / Dz) (grads/period.jl:88-90)
    launch_operator<double>(OP_MKHW, p64.plan, EVv, EVv, Dd.as<c64>(), sel1, dk.as<double>(), hdots.as<double>(), dps,
                            s);
    OMX_HIP_CHECK(hipMemcpyAsync(hh.data(), hdots.as<void>(), sizeof(double) * hh.size(), hipMemcpyDeviceToHost, s));
    OMX_HIP_CHECK(hipStreamSynchronize(s));
    for (int i = 0; i < nb; ++i) dL[i] = 2.0 * sum_part(hh.data() + size_t(i) * dps, dps);
  }

  VecOps V(*this, s, nb, ld, K, EV.as<c64>());
  if (any_ev) {
    OMX_HIP_CHECK(hipMemcpyAsync(actd.as<void>(), hasv.data(), sizeof(int) * nb, hipMemcpyHostToDevice, s));
    const int* act = actd.as<int>();
    for (int i = 0; i < nb; ++i)
      if (hasv[i])
        OMX_HIP_CHECK(hipMemcpyAsync(EB.as<c64>() + i * ld, evbar + 2 * size_t(i) * K, sizeof(c64) * K,
                                     hipMemcpyHostToDevice, s));
    // lambda: (M - w^2) lambda = evbar - ev <ev, evbar> on the complement of ev
    last_adjoint_iters = minres_complement(*this, V, std::vector<double>(nb, omega * omega), dk.as<double>(),
                                           EB.as<c64>(), LA.as<c64>(), hasv, 1e-10);
    // d = fft(kx_tc(ev))/N, ld = fft(kx_tc(lambda)); epsbar += epsinv_bar(d, ld)
    BView<double> LAv{LA.as<c64>(), ld, ld};
    launch_fields<double>(FIELD_D, p64.plan, EVv, Dd.as<c64>(), sel1, dk.as<double>(), 1.0 / double(N), s);
    launch_fields<double>(FIELD_D, p64.plan, LAv, Dl.as<c64>(), sel1, dk.as<double>(), 1.0, s);
    hipLaunchKernelGGL(k_epsbar, dim3(256), dim3(AB), 0, s, Dd.as<c64>(), Dl.as<c64>(), n3, N, nb, act, nullptr, 0,
                       eb6.as<double>());
    // le = bfft(eps^-1 ld / N), e = bfft(eps^-1 d)  (in place)
    launch_fields<double>(FIELD_EINV, p64.plan, LAv, Dl.as<c64>(), sel1, dk.as<double>(), 1.0 / double(N), s);
    launch_fields<double>(FIELD_EINV, p64.plan, EVv, Dd.as<c64>(), sel1, dk.as<double>(), 1.0, s);
    hipLaunchKernelGGL(k_khbar, dim3(nblk, nb), dim3(AB), 0, s, Dl.as<c64>(), Dd.as<c64>(), EV.as<c64>(), LA.as<c64>(),
                       n3, ld, nx, ny, nz, 1.0 / Dx, 1.0 / Dy, nz > 1 ? 1.0 / Dz : 0.0, dk.as<double>(), act,
                       part.as<double>(), Lbar ? partL.as<double>() : nullptr);
    std::vector<double> hp(size_t(nb) * nblk), hl(size_t(nb) * nblk);
    OMX_HIP_CHECK(hipMemcpyAsync(hp.data(), part.as<void>(), sizeof(double) * hp.size(), hipMemcpyDeviceToHost, s));
    if (Lbar)
      OMX_HIP_CHECK(hipMemcpyAsync(hl.data(), partL.as<void>(), sizeof(double) * hl.size(), hipMemcpyDeviceToHost, s));
    OMX_HIP_CHECK(hipStreamSynchronize(s));
    for (int i = 0; i < nb; ++i)
      if (hasv[i]) {
        khb[i] = sum_part(hp.data() + size_t(i) * nblk, nblk);
        if (Lbar) LHb[i] = sum_part(hl.data() + size_t(i) * nblk, nblk);
      }
  }
  // eigenvalue channel: lambda = ((kbar + khbar)/slope) ev -> epsbar += epsinv_bar(d, s N d)
  // and, on 3D cells, Λ̄ += Λ̄_H + k̄tot (-∂ω²∂Λ / ∂ω²∂k) (grads/period.jl:224-232)
  std::vector<double> cf(nb);
  double wb = 0, lb = 0;
  for (int i = 0; i < nb; ++i) {
    const double kb = kbar ? kbar[i] : 0.0;
    const double sc = (kb + khb[i]) / slope[i];
    cf[i] = sc * double(N);
    wb += 2.0 * omega * sc;
    lb += LHb[i] - sc * dL[i];
  }
  OMX_HIP_CHECK(hipMemcpyAsync(cfac.as<void>(), cf.data(), sizeof(double) * nb, hipMemcpyHostToDevice, s));
  OMX_HIP_CHECK(hipMemcpyAsync(actd.as<void>(), all.data(), sizeof(int) * nb, hipMemcpyHostToDevice, s));
  launch_fields<double>(FIELD_D, p64.plan, EVv, Dd.as<c64>(), sel1, dk.as<double>(), 1.0 / double(N), s);
  hipLaunchKernelGGL(k_epsbar, dim3(256), dim3(AB), 0, s, Dd.as<c64>(), Dd.as<c64>(), n3, N, nb, actd.as<int>(),
                     cfac.as<double>(), 1, eb6.as<double>());
  OMX_HIP_CHECK(hipGetLastError());
  std::vector<double> h6(6 * size_t(N));
  OMX_HIP_CHECK(hipMemcpyAsync(h6.data(), eb6.as<void>(), sizeof(double) * h6.size(), hipMemcpyDeviceToHost, s));
  OMX_HIP_CHECK(hipStreamSynchronize(s));
  // (3,3,Nx,Ny) column-major, mirrored (2D convention, grads/solve.jl:490-515); on 3D cells
  // (3,3,Nx,Ny,Nz) with the off-diagonal sums above the diagonal and zeros below (:464-487)
  const bool mirror = nz <= 1;
  for (long p = 0; p < N; ++p) {
    const double* a = h6.data() + 6 * p;
    double* o = epsinvbar + 9 * p;
    o[0] = a[0];
    o[4] = a[1];
    o[8] = a[2];
    o[3] = a[3];
    o[6] = a[4];
    o[7] = a[5];
    o[1] = mirror ? a[3] : 0.0;
    o[2] = mirror ? a[4] : 0.0;
    o[5] = mirror ? a[5] : 0.0;
  }
  *omegabar = wb;
  if (Lbar) *Lbar = lb;
}

// ng_gvd / ng_gvd_E (lib/ModeAnalysis/src/group_index.jl:126-162, :55-95) per mode, FP64:
// the group index n_g = dk/dw from the mode field and GVD = dn_g/dw from one linear solve.
// The reference propagates the sensitivity of n_g backwards (eig_adjt with xbar = dn_g/dev,
// then the basis k-map); here the same derivative is taken forwards, which needs the same
// single solve with M - w^2 and no tensor-valued cotangents.  With Ei = eps^-1, De = deps/dw,
// DDe = d2eps/dw2, B = Ei De Ei, D = fft(kx_tc ev), D_k = fft(d kx_tc/dk ev),
// D_kk = fft(d2 kx_tc/dk2 ev) and Q(X, A, Y) = Re sum_p X_p^H A_p Y_p / N:
//   n_g   = (w + Q(D,B,D)/2) / Q(D,Ei,D_k)                      (= group_index.jl:27)
//   ev'   : (M - w^2) ev' = -n_g dM/dk ev + M_B ev + 2 w ev on the complement of ev, with
//           dM/dk ev = (kx_ct bfft(Ei D_k) - zx_ct bfft(Ei D)) / N
//   num'  = 1 + Q(D,B',D)/2 + Q(D',B,D) + n_g Q(D_k,B,D),  B' = Ei DDe Ei - B De Ei - Ei De B
//   den'  = -Q(D,B,D_k) + Q(D',Ei,D_k) + Q(D,Ei,D'_k) + n_g (Q(D_k,Ei,D_k) + Q(D,Ei,D_kk))
//   GVD   = (num' - n_g den') / Q(D,Ei,D_k)
// E_out (optional): E = Ei . fft(kx_tc(ev)) per mode, (3,Nx,Ny) complex column-major.
void Solver::ng_gvd(double omega, int nev, const double* kmags, const double* evecs, const double* deps,
                    const double* ddeps, double* ng_out, double* gvd_out, double* E_out) {
  require_2d("ng_gvd");
  if (nev < 1 || nev > 64) throw Error(OMX_ERR_UNSUPPORTED, "nev must be in 1..64");
  if (!kmags || !evecs || !deps || !ddeps || !ng_out || !gvd_out) throw Error(OMX_ERR_INVALID_ARG, "NULL argument");
  if (!(omega > 0) || !std::isfinite(omega)) throw Error(OMX_ERR_INVALID_ARG, "omega must be positive and finite");
  for (int i = 0; i < nev; ++i)
    if (!(kmags[i] > 0) || !std::isfinite(kmags[i])) throw Error(OMX_ERR_INVALID_ARG, "kmags must be positive");
  check_symmetric(N, deps, "deps_domega");
  check_symmetric(N, ddeps, "d2eps_domega2");
  OMX_HIP_CHECK(hipSetDevice(device));
  hipStream_t s = stream;
  const int nb = nev;
  const long n3 = 3 * N;
  const double* Ei = epsinv_host.data();
  // B = Ei De Ei and B' = Ei DDe Ei - B De Ei - Ei De B (host, per pixel)
  std::vector<double> B(9 * size_t(N)), Bp(9 * size_t(N)), t1(9 * size_t(N)), t2(9 * size_t(N));
  pixel_product3(N, Ei, deps, Ei, B.data());
  pixel_product3(N, Ei, ddeps, Ei, Bp.data());
  // B De Ei and Ei De B are transposes of each other (all factors symmetric): their sum is
  // twice the symmetric part of B De Ei
  pixel_product3(N, B.data(), deps, Ei, t1.data());
  for (size_t q = 0; q < Bp.size(); ++q) Bp[q] -= 2.0 * t1[q];
  DevMem b6, b6c, bp6, bp6c;
  upload_tensor6<double>(*this, B.data(), b6, b6c);
  upload_tensor6<double>(*this, Bp.data(), bp6, bp6c);
  const double* ei6 = p64.plan.eps6;
  OpPlan<double> planB = p64.plan;
  planB.eps6 = b6.as<double>();
  planB.eps6c = b6c.as<double>();

  const int nblk = int(std::min<long>(256, (N + AB - 1) / AB));
  DevMem dk(sizeof(double) * nb), EV(sizeof(c64) * nb * ld), Hp(sizeof(c64) * nb * ld), R(sizeof(c64) * nb * ld),
      Y(sizeof(c64) * nb * ld), D(sizeof(c64) * nb * n3), Dk(sizeof(c64) * nb * n3), F1(sizeof(c64) * nb * n3),
      F2(sizeof(c64) * nb * n3), qp(sizeof(double) * 9 * nb * nblk), cf(sizeof(double) * 4 * nb);
  OMX_HIP_CHECK(hipMemcpyAsync(dk.as<void>(), kmags, sizeof(double) * nb, hipMemcpyHostToDevice, s));
  OMX_HIP_CHECK(hipMemsetAsync(EV.as<void>(), 0, sizeof(c64) * nb * ld, s));
  for (int i = 0; i < nb; ++i)
    OMX_HIP_CHECK(hipMemcpyAsync(EV.as<c64>() + i * ld, evecs + 2 * size_t(i) * K, sizeof(c64) * K,
                                 hipMemcpyHostToDevice, s));
  const VecSel sel{nb, 1, nullptr, nullptr};
  BView<double> EVv{EV.as<c64>(), ld, ld}, Hpv{Hp.as<c64>(), ld, ld}, Yv{Y.as<c64>(), ld, ld};
  auto qform = [&](int slot, const DevMem& X, const DevMem& Yf, const double* A6) {
    hipLaunchKernelGGL(k_qform, dim3(nblk, nb), dim3(AB), 0, s, X.as<c64>(), Yf.as<c64>(), A6, n3, N,
                       qp.as<double>() + size_t(slot) * nb * nblk);
  };
  auto fetch_q = [&](int slots) {
    std::vector<double> h(size_t(slots) * nb * nblk);
    OMX_HIP_CHECK(hipMemcpyAsync(h.data(), qp.as<void>(), sizeof(double) * h.size(), hipMemcpyDeviceToHost, s));
    OMX_HIP_CHECK(hipStreamSynchronize(s));
    std::vector<double> q(size_t(slots) * nb);
    for (int j = 0; j < slots; ++j)
      for (int i = 0; i < nb; ++i) q[size_t(j) * nb + i] = sum_part(h.data() + (size_t(j) * nb + i) * nblk, nblk) / double(N);
    return q;
  };
  // fields of ev: D, D_k, and D_kk (in F1)
  launch_fields<double>(FIELD_D, p64.plan, EVv, D.as<c64>(), sel, dk.as<double>(), 1.0, s);
  launch_fields<double>(FIELD_DK, p64.plan, EVv, Dk.as<c64>(), sel, dk.as<double>(), 1.0, s);
  launch_fields<double>(FIELD_DKK, p64.plan, EVv, F1.as<c64>(), sel, dk.as<double>(), 1.0, s);
  qform(0, D, D, b6.as<double>());    // Q(D,B,D)
  qform(1, D, Dk, ei6);               // Q(D,Ei,D_k)
  qform(2, D, D, bp6.as<double>());   // Q(D,B',D)
  qform(3, Dk, D, b6.as<double>());   // Q(D_k,B,D)
  qform(4, Dk, Dk, ei6);              // Q(D_k,Ei,D_k)
  qform(5, D, F1, ei6);               // Q(D,Ei,D_kk)
  const std::vector<double> q0 = fetch_q(6);
  std::vector<double> ng(nb);
  std::vector<double> hcf(4 * size_t(nb));
  for (int i = 0; i < nb; ++i) {
    ng[i] = (omega + 0.5 * q0[i]) / q0[size_t(nb) + i];
    hcf[4 * i] = -ng[i] / double(N);
    hcf[4 * i + 1] = ng[i] / double(N);
    hcf[4 * i + 2] = 1.0;
    hcf[4 * i + 3] = 2.0 * omega;
  }
  // right-hand side r = -n_g dM/dk ev + M_B ev + 2 w ev
  OMX_HIP_CHECK(hipMemcpyAsync(F1.as<void>(), Dk.as<void>(), sizeof(c64) * nb * n3, hipMemcpyDeviceToDevice, s));
  OMX_HIP_CHECK(hipMemcpyAsync(F2.as<void>(), D.as<void>(), sizeof(c64) * nb * n3, hipMemcpyDeviceToDevice, s));
  launch_fields<double>(FIELD_EINV, p64.plan, EVv, F1.as<c64>(), sel, dk.as<double>(), 1.0, s);
  launch_fields<double>(FIELD_EINV, p64.plan, EVv, F2.as<c64>(), sel, dk.as<double>(), 1.0, s);
  {
    DevMem Dm(sizeof(c64) * nb * n3);
    launch_operator<double>(OP_M, planB, EVv, Yv, Dm.as<c64>(), sel, dk.as<double>(), nullptr, 0, s);
  }
  OMX_HIP_CHECK(hipMemcpyAsync(cf.as<void>(), hcf.data(), sizeof(double) * hcf.size(), hipMemcpyHostToDevice, s));
  OMX_HIP_CHECK(hipMemsetAsync(R.as<void>(), 0, sizeof(c64) * nb * ld, s));
  hipLaunchKernelGGL(k_ct_combine, dim3(nblk, nb), dim3(AB), 0, s, F1.as<c64>(), F2.as<c64>(), Y.as<c64>(),
                     EV.as<c64>(), n3, ld, nx, ny, 1.0 / Dx, 1.0 / Dy, dk.as<double>(), cf.as<double>(), R.as<c64>());
  OMX_HIP_CHECK(hipGetLastError());
  // ev' on the complement of ev
  VecOps V(*this, s, nb, ld, K, EV.as<c64>());
  const std::vector<int> all(nb, 1);
  last_adjoint_iters = minres_complement(*this, V, std::vector<double>(nb, omega * omega), dk.as<double>(),
                                         R.as<c64>(), Hp.as<c64>(), all, 1e-12);
  // fields of ev': D' (F1) and D'_k (F2)
  launch_fields<double>(FIELD_D, p64.plan, Hpv, F1.as<c64>(), sel, dk.as<double>(), 1.0, s);
  launch_fields<double>(FIELD_DK, p64.plan, Hpv, F2.as<c64>(), sel, dk.as<double>(), 1.0, s);
  qform(0, F1, D, b6.as<double>());   // Q(D',B,D)
  qform(1, F1, Dk, ei6);              // Q(D',Ei,D_k)
  qform(2, D, F2, ei6);               // Q(D,Ei,D'_k)
  const std::vector<double> q1 = fetch_q(3);
  for (int i = 0; i < nb; ++i) {
    const double QDEDk = q0[nb + i], QDBpD = q0[2 * nb + i], QDkBD = q0[3 * nb + i], QDkEDk = q0[4 * nb + i],
                 QDEDkk = q0[5 * nb + i];
    const double QDpBD = q1[i], QDpEDk = q1[nb + i], QDEDpk = q1[2 * nb + i];
    const double g = ng[i];
    const double nump = 1.0 + 0.5 * QDBpD + QDpBD + g * QDkBD;
    const double denp = -QDkBD + QDpEDk + QDEDpk + g * (QDkEDk + QDEDkk);
    ng_out[i] = g;
    gvd_out[i] = (nump - g * denp) / QDEDk;
  }
  if (E_out) {
    launch_fields<double>(FIELD_E, p64.plan, EVv, F1.as<c64>(), sel, dk.as<double>(), 1.0, s);
    OMX_HIP_CHECK(hipMemcpyAsync(E_out, F1.as<void>(), sizeof(c64) * nb * n3, hipMemcpyDeviceToHost, s));
    OMX_HIP_CHECK(hipStreamSynchronize(s));
  }
}

}  // namespace omx
