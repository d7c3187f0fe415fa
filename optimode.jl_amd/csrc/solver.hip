// Host-side engine: plans, batched LOBPCG (DFTK algorithm, lib/MaxwellEigenmodes/src/
// solvers/dftk.jl:304-506) over device-resident blocks, Newton solve_k (src/solve.jl:77-115,
// solvers/gpu.jl:213-242) in an FP32 phase followed by an FP64 refinement phase.
#include "solver.hpp"
#include "krylov.hpp"
#include "refine.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <type_traits>

namespace omx {

hipEvent_t Profiler::take() {
  if (!pool_.empty()) {
    hipEvent_t e = pool_.back();
    pool_.pop_back();
    return e;
  }
  hipEvent_t e;
  OMX_HIP_CHECK(hipEventCreate(&e));
  return e;
}

hipEvent_t Profiler::begin(hipStream_t s) {
  hipEvent_t a = take();
  OMX_HIP_CHECK(hipEventRecord(a, s));
  return a;
}

void Profiler::end(hipStream_t s, hipEvent_t a, KernelStat* st, double bytes) {
  hipEvent_t b = take();
  OMX_HIP_CHECK(hipEventRecord(b, s));
  pending_.push_back(Pending{a, b, st, bytes});
}

void Profiler::flush() {
  for (auto& p : pending_) {
    OMX_HIP_CHECK(hipEventSynchronize(p.b));
    float ms = 0;
    OMX_HIP_CHECK(hipEventElapsedTime(&ms, p.a, p.b));
    p.st->ms += ms;
    p.st->launches += 1;
    p.st->bytes += p.bytes;
    pool_.push_back(p.a);
    pool_.push_back(p.b);
  }
  pending_.clear();
}

Profiler::~Profiler() {
  for (auto& p : pending_) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  for (auto e : pool_) (void)hipEventDestroy(e);
}

template <> PrecisionData<float>& Solver::pd<float>() { return p32; }
template <> PrecisionData<double>& Solver::pd<double>() { return p64; }

// A symmetric per-pixel 3x3 field (column-major (3,3,Nx,Ny) or (3,3,Nx,Ny,Nz), FP64 host) in
// the two device layouts the operator reads: 6 unique entries xx, yy, zz, xy, xz, yz per
// pixel, row-major (iz, iy, ix) and column-major per plane (iz, ix, iy).
template <typename T>
void upload_tensor6(const Solver& S, const double* E, DevMem& rowm, DevMem& colm) {
  const long N = S.N;
  std::vector<T> e6(size_t(6) * N);
  for (long p = 0; p < N; ++p) {
    const double* e = E + 9 * p;
    T* o = e6.data() + 6 * p;
    o[0] = T(e[0]);  // xx
    o[1] = T(e[4]);  // yy
    o[2] = T(e[8]);  // zz
    o[3] = T(e[3]);  // (0,1)
    o[4] = T(e[6]);  // (0,2)
    o[5] = T(e[7]);  // (1,2)
  }
  rowm.ensure(sizeof(T) * 6 * N);
  OMX_HIP_CHECK(hipMemcpy(rowm.as<void>(), e6.data(), sizeof(T) * 6 * N, hipMemcpyHostToDevice));
  // column-major copy per plane: pixel (ix, iy, iz) at 6 (iy + ny (ix + nx iz)) (the
  // register path reads it on 2D cells)
  std::vector<T> e6c(size_t(6) * N);
  const long nxy = long(S.nx) * S.ny;
  for (int iz = 0; iz < S.nz; ++iz)
    for (int ix = 0; ix < S.nx; ++ix)
      for (int iy = 0; iy < S.ny; ++iy)
        for (int a = 0; a < 6; ++a)
          e6c[6 * (iy + size_t(S.ny) * ix + nxy * iz) + a] = e6[6 * (ix + size_t(S.nx) * iy + nxy * iz) + a];
  colm.ensure(sizeof(T) * 6 * N);
  OMX_HIP_CHECK(hipMemcpy(colm.as<void>(), e6c.data(), sizeof(T) * 6 * N, hipMemcpyHostToDevice));
}

template void upload_tensor6<float>(const Solver&, const double*, DevMem&, DevMem&);
template void upload_tensor6<double>(const Solver&, const double*, DevMem&, DevMem&);

// the symmetry test set_epsinv applies, for any per-pixel 3x3 field
void check_symmetric(long N, const double* E, const char* what) {
  double amax = 0;
  for (long i = 0; i < 9 * N; ++i) {
    if (!std::isfinite(E[i])) throw Error(OMX_ERR_INVALID_ARG, std::string(what) + " has non-finite entries");
    amax = std::max(amax, std::fabs(E[i]));
  }
  for (long p = 0; p < N; ++p) {
    const double* e = E + 9 * p;
    for (int a = 0; a < 3; ++a)
      for (int b = a + 1; b < 3; ++b)
        if (std::fabs(e[a + 3 * b] - e[b + 3 * a]) > 1e-12 * amax)
          throw Error(OMX_ERR_UNSUPPORTED, std::string(what) + " must be symmetric per pixel");
  }
}

// per pixel, out = sym(L M R) for column-major 3x3 fields (3,3,Nx,Ny)
void pixel_product3(long N, const double* L, const double* M, const double* R, double* out) {
  for (long p = 0; p < N; ++p) {
    const double *l = L + 9 * p, *m = M + 9 * p, *r = R + 9 * p;
    double t[9];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) t[a + 3 * b] = l[a] * m[3 * b] + l[a + 3] * m[1 + 3 * b] + l[a + 6] * m[2 + 3 * b];
    double* o = out + 9 * p;
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) o[a + 3 * b] = t[a] * r[3 * b] + t[a + 3] * r[1 + 3 * b] + t[a + 6] * r[2 + 3 * b];
    for (int a = 0; a < 3; ++a)
      for (int b = a + 1; b < 3; ++b) o[a + 3 * b] = o[b + 3 * a] = 0.5 * (o[a + 3 * b] + o[b + 3 * a]);
  }
}

namespace {

constexpr double EPS32 = 1.1920928955078125e-07;
constexpr double EPS64 = 2.220446049250313e-16;

template <typename T> constexpr double eps_of() { return sizeof(T) == 4 ? EPS32 : EPS64; }

// every radix pass of a batched LDS FFT (nb sequences of f.L) fits `threads` threads
// holding at most vpt values each (the bound fft_pass relies on)
bool fft_fits(int nb, const FFTFactors& f, int threads, int vpt) {
  for (int p = 0; p < f.n; ++p) {
    const int R = f.r[p];
    const long nbf = long(nb) * (f.L / R);
    if ((nbf + threads - 1) / threads > (vpt + R - 1) / R) return false;
  }
  return true;
}

int choose_vpt(int nb, const FFTFactors& f) {
  for (int v : {8, 16, 32}) {
    bool ok = true;
    for (int p = 0; p < f.n; ++p) {
      const int R = f.r[p];
      const int nbf = nb * (f.L / R);
      const int need = (nbf + 255) / 256;
      if (need > (v + R - 1) / R) ok = false;
    }
    if (ok) return v;
  }
  return -1;
}

template <typename T> void select_operator(Solver& S, PrecisionData<T>& P);

template <typename T> void build_plan(Solver& S, PrecisionData<T>& P) {
  OpPlan<T>& pl = P.plan;
  pl.nx = S.nx;
  pl.ny = S.ny;
  pl.invdx = 1.0 / S.Dx;
  pl.invdy = 1.0 / S.Dy;
  if (!fft_factor(S.nx, pl.fx) || !fft_factor(S.ny, pl.fy))
    throw Error(OMX_ERR_BAD_SHAPE, "grid sizes must factor into 2, 3 and 5 (got " + std::to_string(S.nx) + " x " +
                                       std::to_string(S.ny) + ")");
  const bool tuned = sizeof(T) == 4 && S.nz == 1;  // options of the FP32 plan of 2D cells
  if (tuned && S.opt.fft_variant32 == 1) {
    fft_factor_radix4(S.nx, pl.fx);
    fft_factor_radix4(S.ny, pl.fy);
  }
  auto twiddles = [](int L) {
    std::vector<cplx<T>> tw(L);
    for (int t = 0; t < L; ++t) {
      const long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)t / (long double)L;
      tw[t] = cplx<T>{T(std::cos(a)), T(std::sin(a))};
    }
    return tw;
  };
  pl.nz = S.nz;
  pl.invdz = S.nz > 1 ? 1.0 / S.Dz : 0.0;
  pl.twz = nullptr;
  pl.zcols_per_block = 1;
  if (S.nz > 1) {
    if (!fft_factor(S.nz, pl.fz))
      throw Error(OMX_ERR_BAD_SHAPE, "grid sizes must factor into 2, 3 and 5 (got Nz = " + std::to_string(S.nz) + ")");
    auto twz = twiddles(S.nz);
    P.twz.alloc(sizeof(cplx<T>) * S.nz);
    OMX_HIP_CHECK(hipMemcpy(P.twz.template as<void>(), twz.data(), sizeof(cplx<T>) * S.nz, hipMemcpyHostToDevice));
    pl.twz = P.twz.template as<cplx<T>>();
  }
  auto twx = twiddles(S.nx), twy = twiddles(S.ny);
  P.twx.alloc(sizeof(cplx<T>) * S.nx);
  P.twy.alloc(sizeof(cplx<T>) * S.ny);
  OMX_HIP_CHECK(hipMemcpy(P.twx.template as<void>(), twx.data(), sizeof(cplx<T>) * S.nx, hipMemcpyHostToDevice));
  OMX_HIP_CHECK(hipMemcpy(P.twy.template as<void>(), twy.data(), sizeof(cplx<T>) * S.ny, hipMemcpyHostToDevice));
  pl.twx = P.twx.template as<cplx<T>>();
  pl.twy = P.twy.template as<cplx<T>>();
  // rows per block (passes 1/3) and columns per block (pass 2): ~3k complex values in LDS
  const int budget = sizeof(T) == 4 ? 3072 : 1536;
  pl.rows_per_block = 1;
  for (int rb : {8, 4, 2}) {
    if (S.ny % rb == 0 && 3 * rb * S.nx <= budget) {
      pl.rows_per_block = rb;
      break;
    }
  }
  pl.cols_per_block = 1;
  for (int tx : {16, 8, 4, 2}) {
    if (S.nx % tx == 0 && 3 * tx * S.ny <= budget) {
      pl.cols_per_block = tx;
      break;
    }
  }
  // tuned tile shape: any divisor that stays within the LDS budget (the FFT bound is
  // checked below with the rest of the plan)
  if (tuned && S.opt.tile_rows32 > 0 && S.ny % S.opt.tile_rows32 == 0 && 3 * S.opt.tile_rows32 * S.nx <= budget)
    pl.rows_per_block = S.opt.tile_rows32;
  if (tuned && S.opt.tile_cols32 > 0 && S.nx % S.opt.tile_cols32 == 0 && 3 * S.opt.tile_cols32 * S.ny <= budget)
    pl.cols_per_block = S.opt.tile_cols32;
  if (S.nz > 1)
    for (int tz : {16, 8, 4, 2}) {
      if (S.nx % tz == 0 && 3 * tz * S.nz <= budget) {
        pl.zcols_per_block = tz;
        break;
      }
    }
  const int v1 = choose_vpt(3 * pl.rows_per_block, pl.fx);
  const int v2 = choose_vpt(3 * pl.cols_per_block, pl.fy);
  const int v3 = S.nz > 1 ? choose_vpt(3 * pl.zcols_per_block, pl.fz) : 8;
  if (v1 < 0 || v2 < 0 || v3 < 0) throw Error(OMX_ERR_BAD_SHAPE, "grid too large for the LDS FFT");
  pl.vpt = std::max({v1, v2, v3});
  const size_t lds13 = sizeof(cplx<T>) * (size_t(S.nx) + 3 * size_t(pl.rows_per_block) * S.nx);
  const size_t lds2 = sizeof(cplx<T>) * (size_t(S.ny) + 3 * size_t(pl.cols_per_block) * S.ny);
  const size_t lds3 = sizeof(cplx<T>) * (size_t(S.nz) + 3 * size_t(pl.zcols_per_block) * S.nz);
  // yz-slab path: whole x-columns (3 ny nz values) per block when they fit in LDS and in
  // the slab kernel's FFT bound
  P.slab_ok = 0;
  if (S.nz > 1 && sizeof(cplx<T>) * (size_t(S.nx) + 3 * size_t(pl.rows_per_block) * (S.nx + 1)) <= 160 * 1024)
    for (int ts : {2, 1}) {
      const size_t ls = sizeof(cplx<T>) * (size_t(S.ny) + S.nz + 3 * size_t(ts) * S.ny * S.nz);
      if (S.nx % ts == 0 && ls <= 160 * 1024 && fft_fits(3 * ts * S.nz, pl.fy, SLAB_THREADS, SLAB_VPT) &&
          fft_fits(3 * ts * S.ny, pl.fz, SLAB_THREADS, SLAB_VPT)) {
        P.slab_ok = ts;
        break;
      }
    }
  if (lds13 > 160 * 1024 || lds2 > 160 * 1024 || lds3 > 160 * 1024)
    throw Error(OMX_ERR_BAD_SHAPE, "grid too large for the LDS FFT");
  // register-FFT operator for power-of-two sizes 64 .. 512 (2D cells)
  P.reg_ok = (S.nz == 1 && reg_fft_ok(S.nx) && reg_fft_ok(S.ny) && S.ny % reg_rows<T>() == 0 &&
              S.nx % (REG_COL_THREADS / (S.ny / 8)) == 0)
                 ? 1
                 : 0;
  select_operator(S, P);
}

// operator implementation per the grid and the reg_ops option; the OP_MKH dot partial
// count follows it
template <typename T> void select_operator(Solver& S, PrecisionData<T>& P) {
  OpPlan<T>& pl = P.plan;
  pl.reg = (P.reg_ok && S.opt.reg_ops) ? 1 : 0;
  pl.reg_rb = sizeof(T) == 4 ? S.opt.reg_rows32 : S.opt.reg_rows64;
  pl.reg_nt2 = S.opt.reg_nt2;
  pl.slab_cols = S.opt.slab3d ? P.slab_ok : 0;
  pl.dot_blocks = S.nz * (S.ny / (pl.reg ? reg_rows_of(pl) : pl.rows_per_block));
}

template <typename T> void upload_eps(Solver& S, PrecisionData<T>& P) {
  upload_tensor6<T>(S, S.epsinv_host.data(), P.eps6, P.eps6c);
  P.plan.eps6 = P.eps6.template as<T>();
  P.plan.eps6c = P.eps6c.template as<T>();
}

// ---------------------------------------------------------------------------------------
// Batched LOBPCG over ninst instances with block size M each, vectors of precision T.
// ---------------------------------------------------------------------------------------
template <typename T> class Batch {
 public:
  using value_type = T;
  Solver& S;
  int ninst, M;
  long K, ld, istride;
  DevMem X[2], AX[2], R[2], AR, P[2], AP[2], D;
  DevMem st, shape, gramp, red, normp, Cx, Cp, Cw, lam, rhist, dotp, active, band, kd, ns, selc0, selnc;
  DenseLaunch dl{};
  PinnedMem snap[2];  // LState snapshots of consecutive iterations
  hipEvent_t snap_ev[2] = {nullptr, nullptr};
  int pm, nchunk, nblk, qmax, dps;
  int nchunk_max, nblk_max;  // for one active instance (buffer sizes)
  int cur = 0;
  int maxiter_alloc;
  std::vector<LState> hst;
  hipStream_t sm;  // the stream every launch and copy of this batch goes to

  Batch(Solver& s, int ninst_, int M_, int maxiter, hipStream_t stream_ = nullptr)
      : S(s), ninst(ninst_), M(M_), maxiter_alloc(maxiter), sm(stream_ ? stream_ : s.stream) {
    K = S.K;
    ld = S.ld;
    istride = long(M) * ld;
    const size_t blk = sizeof(cplx<T>) * size_t(ninst) * istride;
    for (int b = 0; b < 2; ++b) {
      X[b].alloc(blk);
      AX[b].alloc(blk);
      R[b].alloc(blk);
      P[b].alloc(blk);
      AP[b].alloc(blk);
    }
    AR.alloc(blk);
    D.alloc(sizeof(cplx<T>) * size_t(ninst) * M * 3 * S.N);
    pm = (3 * M <= 16) ? 16 : (3 * M <= 32 ? 32 : 48);
    qmax = M <= 4 ? 4 : (M <= 8 ? 8 : 16);
    // about 512 Gram blocks (2 per CU), at least 512 rows each
    nchunk_max = std::max(1, std::min<int>(int((K + 511) / 512), 512));
    nblk_max = std::max(1, std::min<int>(int((K + 255) / 256), 1024));
    set_active(ninst);
    dps = S.pd<T>().plan.dot_blocks;
    st.alloc(sizeof(LState) * ninst);
    shape.alloc(sizeof(int) * F_NFIELDS * ninst);
    OMX_HIP_CHECK(hipMemset(shape.as<void>(), 0, sizeof(int) * F_NFIELDS * ninst));
    gramp.alloc(sizeof(c64) * size_t(ninst) * nchunk_max * pm * pm);
    red.alloc(sizeof(c64) * size_t(ninst) * pm * pm);
    normp.alloc(sizeof(double) * size_t(ninst) * nblk_max * qmax);
    Cx.alloc(sizeof(c64) * size_t(ninst) * pm * pm);
    Cp.alloc(sizeof(c64) * size_t(ninst) * pm * pm);
    Cw.alloc(sizeof(c64) * size_t(ninst) * pm * pm);
    lam.alloc(sizeof(double) * size_t(ninst) * pm);
    rhist.alloc(sizeof(double) * size_t(ninst) * (maxiter + 1) * 16);
    dotp.alloc(sizeof(double) * size_t(ninst) * M * dps);
    active.alloc(sizeof(int) * ninst);
    band.alloc(sizeof(int) * ninst);
    kd.alloc(sizeof(double) * ninst);
    ns.alloc(sizeof(NewtonState) * ninst);
    selc0.alloc(sizeof(int) * ninst);
    selnc.alloc(sizeof(int) * ninst);
    dl.ninst = ninst;
    dl.pm = pm;
    dl.nchunk = 1;  // dense kernels read the reduced Gram matrices
    dl.qmax = qmax;
    dl.jac_stop = 1e-9;
    hst.resize(ninst);
    for (int b = 0; b < 2; ++b) {
      snap[b].alloc(sizeof(LState) * ninst);
      OMX_HIP_CHECK(hipEventCreateWithFlags(&snap_ev[b], hipEventDisableTiming));
    }
  }
  ~Batch() {
    for (int b = 0; b < 2; ++b)
      if (snap_ev[b]) (void)hipEventDestroy(snap_ev[b]);
  }
  Batch(const Batch&) = delete;
  Batch& operator=(const Batch&) = delete;

  // Row partitioning of the Gram products and block updates for `na` instances with work
  // (an upper bound is enough: the kernels skip instances without columns): about
  // gram_blocks Gram blocks and combine_blocks update blocks in total (by default 1024 and
  // 2048 on 2D cells, 512 and 1024 on 3D cells), whatever the number of instances left.
  void set_active(int na) {
    na = std::max(1, std::min(na, ninst));
    const bool flat = S.nz <= 1;
    const int gb = S.opt.gram_blocks > 0 ? S.opt.gram_blocks : (flat ? 1024 : 512);
    const int cb = S.opt.combine_blocks > 0 ? S.opt.combine_blocks : (flat ? 2048 : 1024);
    nchunk = std::max(1, std::min(nchunk_max, gb / na));
    nblk = std::max(1, std::min(nblk_max, cb / na));
    dl.nblk = nblk;
  }

  cplx<T>* p(DevMem& m) { return m.as<cplx<T>>(); }
  int* f(int field) { return shape.as<int>() + field * ninst; }
  Part<T> part(DevMem& m, const int* off, const int* cnt, int foff = 0, int fcnt = 0) {
    return Part<T>{m.as<cplx<T>>(), ld, istride, ShapeRef{off, cnt, foff, fcnt}};
  }
  OutView<T> out(DevMem& m, const int* off, int foff = 0) { return OutView<T>{m.as<cplx<T>>(), ld, istride, off, foff}; }
  static ColList<T> list(std::initializer_list<Part<T>> ps) {
    ColList<T> L{};
    L.n = 0;
    for (const auto& q : ps) L.p[L.n++] = q;
    return L;
  }
  CombineCoef coef(DevMem& C, const int* qcnt, const double* lamp = nullptr) {
    return CombineCoef{C.as<double>(), long(pm) * pm, pm, qcnt, 0, lamp, pm};
  }

  void apply(OpKind op, DevMem& in, DevMem& outm, const int* col0, const int* ncol, double* dots = nullptr) {
    VecSel sel{ninst, M, col0, ncol};
    BView<T> Xv{p(in), ld, istride}, Yv{p(outm), ld, istride};
    launch_operator<T>(op, S.pd<T>().plan, Xv, Yv, p(D), sel, kd.as<double>(), dots, dps, sm);
  }
  // red_i = A_i^H B_i (pmv x pmv per instance, default pm); ca, cb: upper bounds of the
  // column counts of A and B; bpre: B is the leading columns of A
  void gram(const ColList<T>& A, const ColList<T>& B, int ca, int cb, int pmv = -1, bool bpre = false,
            bool herm = false) {
    const int pg = pmv < 0 ? pm : pmv;
    launch_gram<T>(A, B, ninst, K, nchunk, pg, gramp.as<c64>(), sm, ca, cb, bpre, herm);
    // instances whose A has no columns were skipped by the Gram: skip their reduction too
    launch_reduce(gramp.as<c64>(), ninst, nchunk, pg * pg, red.as<c64>(), sm, A.p[0].s.cnt);
  }
  // R <- R U^-1 twice (CholeskyQR2), R = list of one part with offset/count fields
  void cholqr(DevMem& Rm, const int* off, int cnt_field) {
    ColList<T> Rl = list({part(Rm, off, f(cnt_field))});
    ColList<T> none{};
    for (int pass = 0; pass < 2; ++pass) {
      gram(Rl, Rl, M, M, -1, true);
      launch_dense_chol(dl, st.as<LState>(), shape.as<int>(), cnt_field, red.as<c64>(), Cw.as<double>(), sm);
      launch_combine<T>(CMB_PLAIN, Rl, none, coef(Cw, f(cnt_field)), qmax, out(Rm, off), out(Rm, off), out(Rm, off),
                        ninst, K, nblk, nullptr, sm);
    }
  }
  // ortho!(R, [X P]) (dftk.jl:487-495) as two passes of R <- (R - Z Z^H R) U^{-1}, each
  // one Gram of [R Z]^H R, one small dense step and one in-place update; R = F_NNEW
  // columns at offset F_NEWLY, Z = X_n (M columns) and P_n (F_NPN columns).  R leads the
  // list so that the Gram reads its columns once.
  void ortho_r(DevMem& Xn, DevMem& Pn, DevMem& Rm) {
    ColList<T> Rl = list({part(Rm, f(F_NEWLY), f(F_NNEW))});
    ColList<T> RZ = list({part(Rm, f(F_NEWLY), f(F_NNEW)), part(Xn, nullptr, f(F_M)), part(Pn, nullptr, f(F_NPN))});
    ColList<T> none{};
    for (int pass = 0; pass < 2; ++pass) {
      gram(RZ, Rl, 3 * M, M, -1, M <= 16);
      launch_dense_orth(dl, st.as<LState>(), shape.as<int>(), red.as<c64>(), pm, Cw.as<double>(), sm);
      launch_combine<T>(CMB_PLAIN, RZ, none, coef(Cw, f(F_NNEW)), qmax, out(Rm, f(F_NEWLY)), out(Rm, f(F_NEWLY)),
                        out(Rm, f(F_NEWLY)), ninst, K, nblk, nullptr, sm);
    }
  }

  // Run LOBPCG on every active instance until all finished.  Returns iterations summed
  // over instances; per-instance status in hst.
  // msize[i]: block size of instance i (<= M); tols[i]: its locking threshold.
  // The host reads the per-instance state one iteration late: iteration it is enqueued
  // before the state after iteration it - 1 is examined, so the device never waits for the
  // host.  An iteration enqueued after every instance has finished does nothing (every
  // kernel skips instances with no active columns).  run_begin / run_step / run_end are the
  // pieces of run, so that two batches on two streams can advance in lockstep (run2).
  struct OpMark {
    hipEvent_t a, b;
  };
  struct RunState {
    std::vector<int> act, nact;  // nact: active columns entering the next examined iteration
    std::vector<LState> init;     // host source of the initial states (kept until the copy ran)
    std::vector<OpMark> marks;    // operator intervals, bytes attributed on examination
    int it = 0;
    int maxiter = 0;
    bool done = false;
    double host_enqueue_ms = 0, host_wait_ms = 0;  // host time launching / waiting (OMX_TRACE)
  };
  void run_begin(RunState& r, const std::vector<int>& act, const std::vector<int>& bands,
                 const std::vector<int>& ncheck, const std::vector<int>& msize, const std::vector<double>& tols,
                 int maxiter, SolveStats& stats) {
    if (maxiter > maxiter_alloc) maxiter = maxiter_alloc;
    r = RunState{};
    r.act = act;
    r.maxiter = maxiter;
    dl.maxiter = maxiter;
    dl.jac_stop = sizeof(T) == 4 ? S.opt.jac_stop32 : 1e-9;
    r.init.assign(ninst, LState{});
    for (int i = 0; i < ninst; ++i) {
      LState& s = r.init[i];
      std::memset(&s, 0, sizeof(s));
      s.M = msize[i];
      s.ncheck = ncheck[i];
      s.tol = tols[i];
      s.eps = eps_of<T>();
      s.ortho_tol = 2 * eps_of<T>();
    }
    OMX_HIP_CHECK(hipMemcpyAsync(st.as<void>(), r.init.data(), sizeof(LState) * ninst, hipMemcpyHostToDevice, sm));
    OMX_HIP_CHECK(hipMemcpyAsync(active.as<void>(), act.data(), sizeof(int) * ninst, hipMemcpyHostToDevice, sm));
    OMX_HIP_CHECK(hipMemcpyAsync(band.as<void>(), bands.data(), sizeof(int) * ninst, hipMemcpyHostToDevice, sm));
    launch_lobpcg_init(dl, st.as<LState>(), shape.as<int>(), active.as<int>(), band.as<int>(), sm);
    long ncols_active = 0;
    for (int i = 0; i < ninst; ++i) ncols_active += act[i] ? msize[i] : 0;
    // X = ortho!(X); AX = A X   (dftk.jl:317, :326)
    cholqr(X[cur], nullptr, F_M);
    apply(OP_M, X[cur], AX[cur], f(F_XOFF), f(F_M));
    stats.matvec += ncols_active;
    r.nact.assign(ninst, 0);
    for (int i = 0; i < ninst; ++i) r.nact[i] = act[i] ? msize[i] : 0;
    int na = 0;
    for (int i = 0; i < ninst; ++i) na += act[i] ? 1 : 0;
    set_active(na);
    r.it = 0;
    r.done = false;
  }

  // Enqueue iteration r.it, then (from the second on) examine the state after the one
  // before it.  Sets r.done once every instance has finished.
  void run_step(RunState& r, SolveStats& stats) {
    using clk = std::chrono::steady_clock;
    const auto t_enq = clk::now();
    const int it = r.it;
    const bool prof = S.prof.enabled;
    const int c = cur, n = cur ^ 1;
    if (it > 0) {
      hipEvent_t a = prof ? S.prof.begin(sm) : nullptr;
      apply(OP_M, R[c], AR, f(F_ROFF), f(F_NACT));
      if (prof) r.marks.push_back(OpMark{a, S.prof.begin(sm)});
    }
    ColList<T> Y, AY;
    if (it == 0) {
      Y = list({part(X[c], f(F_XOFF), f(F_NACT))});
      AY = list({part(AX[c], f(F_XOFF), f(F_NACT))});
    } else {
      Y = list({part(X[c], f(F_XOFF), f(F_NACT)), part(R[c], f(F_ROFF), f(F_NACT)), part(P[c], nullptr, f(F_NP))});
      AY = list({part(AX[c], f(F_XOFF), f(F_NACT)), part(AR, f(F_ROFF), f(F_NACT)), part(AP[c], nullptr, f(F_NP))});
    }
    // Y^H A Y is Hermitian: upper tiles only (the Rayleigh-Ritz step mirrors them)
    gram(Y, AY, it == 0 ? M : 3 * M, it == 0 ? M : 3 * M, -1, false, true);
    launch_dense_rr(dl, st.as<LState>(), shape.as<int>(), red.as<c64>(), Cx.as<double>(), lam.as<double>(),
                    it == 0 ? 1 : 0, sm);
    launch_combine<T>(CMB_RESID, Y, AY, coef(Cx, f(F_NACT), lam.as<double>()), qmax, out(X[n], f(F_XOFF)),
                      out(AX[n], f(F_XOFF)), out(R[n], nullptr, 0), ninst, K, nblk, normp.as<double>(), sm);
    launch_dense_lock(dl, st.as<LState>(), shape.as<int>(), normp.as<double>(), Cx.as<double>(), Cp.as<double>(),
                      rhist.as<double>(), sm);
    apply(OP_P, R[n], R[n], f(F_NEWLY), f(F_NNEW));
    if (it > 0)
      launch_combine<T>(CMB_TWO, Y, AY, coef(Cp, f(F_NNEW)), qmax, out(P[n], nullptr, 0), out(AP[n], nullptr, 0),
                        out(AP[n], nullptr, 0), ninst, K, nblk, nullptr, sm);
    launch_copy_cols<T>(p(X[n]), p(X[c]), ld, istride, ninst, f(F_XOFF), f(F_NCOPY), K, sm, p(AX[n]), p(AX[c]));
    ortho_r(X[n], P[n], R[n]);
    launch_advance(dl, st.as<LState>(), shape.as<int>(), sm);
    cur = n;
    OMX_HIP_CHECK(hipMemcpyAsync(snap[it & 1].as<void>(), st.as<void>(), sizeof(LState) * ninst,
                                 hipMemcpyDeviceToHost, sm));
    OMX_HIP_CHECK(hipEventRecord(snap_ev[it & 1], sm));
    r.it = it + 1;
    const auto t_wait = clk::now();
    r.host_enqueue_ms += std::chrono::duration<double, std::milli>(t_wait - t_enq).count();
    if (it == 0) return;
    // examine the state after iteration it - 1
    OMX_HIP_CHECK(hipEventSynchronize(snap_ev[(it - 1) & 1]));
    r.host_wait_ms += std::chrono::duration<double, std::milli>(clk::now() - t_wait).count();
    const LState* hs = snap[(it - 1) & 1].as<LState>();
    bool all = true;
    long m_next = 0;
    for (int i = 0; i < ninst; ++i) {
      if (!r.act[i]) continue;
      if (r.nact[i] > 0) {
        stats.precon += r.nact[i];
        if (it - 1 > 0) stats.matvec += r.nact[i];
      }
      if (!hs[i].finished) all = false;
      r.nact[i] = hs[i].finished ? 0 : hs[i].nact;
      m_next += r.nact[i];
    }
    {
      int na = 0;
      for (int i = 0; i < ninst; ++i) na += r.nact[i] > 0 ? 1 : 0;
      set_active(na);
    }
    if (prof) {
      // iteration it applied the operator to the columns active after iteration it - 1
      const double w = double(sizeof(T)) / 4.0;
      const OpMark& mk = r.marks.back();
      if (m_next > 0)
        S.prof.commit(mk.a, mk.b, sizeof(T) == 4 ? &S.prof.op32 : &S.prof.op64,
                      w * (32.0 * S.N * m_next + 24.0 * S.N));
      else
        S.prof.recycle(mk.a, mk.b);
    }
    if (all) {
      r.done = true;
      return;
    }
    if (it > r.maxiter + 2) throw Error(OMX_ERR_INTERNAL, "LOBPCG driver did not terminate");
  }

  // The final states (the iteration enqueued last changed nothing); returns iterations
  // summed over the active instances.
  int run_end(RunState& r) {
    const int last = r.it - 1;
    OMX_HIP_CHECK(hipEventSynchronize(snap_ev[last & 1]));
    std::memcpy(hst.data(), snap[last & 1].as<LState>(), sizeof(LState) * ninst);
    int tot = 0;
    for (int i = 0; i < ninst; ++i)
      if (r.act[i]) tot += hst[i].niter;
    if (S.trace)
      std::fprintf(stderr, "[omx] %s lobpcg run: %d rounds, host enqueue %.2f ms, host wait %.2f ms\n",
                   sizeof(T) == 4 ? "fp32" : "fp64", r.it, r.host_enqueue_ms, r.host_wait_ms);
    for (int i = 0; i < ninst; ++i)
      if (r.act[i] && hst[i].status == OMX_ERR_CHOLESKY)
        throw Error(OMX_ERR_CHOLESKY, "Cholesky orthogonalisation failed even after shifting");
    return tot;
  }

  int run(const std::vector<int>& act, const std::vector<int>& bands, const std::vector<int>& ncheck,
          const std::vector<int>& msize, const std::vector<double>& tols, int maxiter, SolveStats& stats) {
    RunState r;
    run_begin(r, act, bands, ncheck, msize, tols, maxiter, stats);
    while (!r.done) run_step(r, stats);
    return run_end(r);
  }

  // One Newton update for every active instance from its band's Rayleigh quotient and HMkH.
  void newton(std::vector<NewtonState>& hns) {
    OMX_HIP_CHECK(hipMemcpyAsync(ns.as<void>(), hns.data(), sizeof(NewtonState) * ninst, hipMemcpyHostToDevice, sm));
    ColList<T> a = list({part(X[cur], f(F_EIG), f(F_ONE))});
    ColList<T> b = list({part(AX[cur], f(F_EIG), f(F_ONE))});
    gram(a, b, 1, 1, 16);
    apply(OP_MKH, X[cur], X[cur], f(F_EIG), f(F_ONE), dotp.as<double>());
    launch_newton(ninst, M, ns.as<NewtonState>(), st.as<LState>(), red.as<c64>(), 1, 16, dotp.as<double>(), dps,
                  kd.as<double>(), sm);
    OMX_HIP_CHECK(hipMemcpyAsync(hns.data(), ns.as<void>(), sizeof(NewtonState) * ninst, hipMemcpyDeviceToHost, sm));
    OMX_HIP_CHECK(hipStreamSynchronize(sm));
  }

  // HMkH of the first ncol columns of instance `inst` (others idle), summed on the host.
  std::vector<double> hmkh_cols(int inst, int ncol) {
    std::vector<int> c0(ninst, 0), nc(ninst, 0);
    nc[inst] = ncol;
    DevMem &dc0 = selc0, &dnc = selnc;
    OMX_HIP_CHECK(hipMemcpyAsync(dc0.as<void>(), c0.data(), sizeof(int) * ninst, hipMemcpyHostToDevice, sm));
    OMX_HIP_CHECK(hipMemcpyAsync(dnc.as<void>(), nc.data(), sizeof(int) * ninst, hipMemcpyHostToDevice, sm));
    apply(OP_MKH, X[cur], X[cur], dc0.as<int>(), dnc.as<int>(), dotp.as<double>());
    std::vector<double> h(size_t(ninst) * M * dps);
    OMX_HIP_CHECK(hipMemcpyAsync(h.data(), dotp.as<void>(), sizeof(double) * h.size(), hipMemcpyDeviceToHost, sm));
    OMX_HIP_CHECK(hipStreamSynchronize(sm));
    std::vector<double> out(ncol);
    for (int j = 0; j < ncol; ++j) {
      double acc = 0;
      for (int b = 0; b < dps; ++b) acc += h[(size_t(inst) * M + j) * dps + b];
      out[j] = acc;
    }
    return out;
  }

  // HMkH of the first ncol[i] columns of every instance i (out[i * M + j]), summed on the host.
  std::vector<double> hmkh_multi(const std::vector<int>& ncol) {
    std::vector<int> c0(ninst, 0);
    OMX_HIP_CHECK(hipMemcpyAsync(selc0.as<void>(), c0.data(), sizeof(int) * ninst, hipMemcpyHostToDevice, sm));
    OMX_HIP_CHECK(hipMemcpyAsync(selnc.as<void>(), ncol.data(), sizeof(int) * ninst, hipMemcpyHostToDevice, sm));
    apply(OP_MKH, X[cur], X[cur], selc0.as<int>(), selnc.as<int>(), dotp.as<double>());
    std::vector<double> h(size_t(ninst) * M * dps);
    OMX_HIP_CHECK(hipMemcpyAsync(h.data(), dotp.as<void>(), sizeof(double) * h.size(), hipMemcpyDeviceToHost, sm));
    OMX_HIP_CHECK(hipStreamSynchronize(sm));
    std::vector<double> out(size_t(ninst) * M, 0.0);
    for (int i = 0; i < ninst; ++i)
      for (int j = 0; j < ncol[i]; ++j) {
        double acc = 0;
        for (int q = 0; q < dps; ++q) acc += h[(size_t(i) * M + j) * dps + q];
        out[size_t(i) * M + j] = acc;
      }
    return out;
  }

  // Rayleigh quotients diag(X^H A X) of all M columns of every instance (final_retval).
  std::vector<double> rayleigh() {
    ColList<T> a = list({part(X[cur], nullptr, nullptr, 0, M)});
    ColList<T> b = list({part(AX[cur], nullptr, nullptr, 0, M)});
    gram(a, b, M, M);
    std::vector<c64> h(size_t(ninst) * pm * pm);
    OMX_HIP_CHECK(hipMemcpyAsync(h.data(), red.as<void>(), sizeof(c64) * h.size(), hipMemcpyDeviceToHost, sm));
    OMX_HIP_CHECK(hipStreamSynchronize(sm));
    std::vector<double> out(size_t(ninst) * M);
    for (int i = 0; i < ninst; ++i)
      for (int j = 0; j < M; ++j) out[size_t(i) * M + j] = h[size_t(i) * pm * pm + j * pm + j].x;
    return out;
  }
};

}  // namespace

// ---------------------------------------------------------------------------------------
// Solver
// ---------------------------------------------------------------------------------------

Solver::Solver(int dev, int nx_, int ny_, double Dx_, double Dy_, const double* epsinv)
    : Solver(dev, nx_, ny_, 1, Dx_, Dy_, 0.0, epsinv) {}

Solver::Solver(int dev, int nx_, int ny_, int nz_, double Dx_, double Dy_, double Dz_, const double* epsinv)
    : device(dev), nx(nx_), ny(ny_), Dx(Dx_), Dy(Dy_), nz(nz_), Dz(Dz_) {
  const char* tr = std::getenv("OMX_TRACE");
  trace = tr && tr[0] && tr[0] != '0';
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) throw Error(OMX_ERR_NO_DEVICE, "no HIP device available");
  if (dev < 0 || dev >= count) throw Error(OMX_ERR_NO_DEVICE, "device index out of range");
  hipDeviceProp_t prop;
  OMX_HIP_CHECK(hipGetDeviceProperties(&prop, dev));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    throw Error(OMX_ERR_NO_DEVICE, std::string("omx is built for gfx950, device is ") + prop.gcnArchName);
  if (nx < 2 || ny < 2) throw Error(OMX_ERR_BAD_SHAPE, "grid must be at least 2 x 2");
  if (nz < 1) throw Error(OMX_ERR_BAD_SHAPE, "Nz must be positive");
  if (!(Dx > 0) || !(Dy > 0) || !std::isfinite(Dx) || !std::isfinite(Dy))
    throw Error(OMX_ERR_INVALID_ARG, "cell size must be positive and finite");
  if (nz > 1 && !(Dz > 0 && std::isfinite(Dz))) throw Error(OMX_ERR_INVALID_ARG, "the period Dz must be positive");
  OMX_HIP_CHECK(hipSetDevice(dev));
  noalloc_init(dev);
  OMX_HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  OMX_HIP_CHECK(hipStreamCreateWithFlags(&stream2, hipStreamNonBlocking));
  N = long(nx) * ny * nz;
  K = 2 * N;
  ld = (K + 63) / 64 * 64;
  build_plan<float>(*this, p32);
  build_plan<double>(*this, p64);
  set_epsinv(epsinv);
}

Solver::~Solver() {
  workspaces.clear();  // batches hold events on the streams
  if (stream2) {
    (void)hipStreamSynchronize(stream2);
    (void)hipStreamDestroy(stream2);
  }
  if (stream) {
    (void)hipStreamSynchronize(stream);
    (void)hipStreamDestroy(stream);
  }
}

void Solver::set_epsinv(const double* epsinv) {
  if (!epsinv) throw Error(OMX_ERR_INVALID_ARG, "epsinv is NULL");
  const long n9 = 9 * N;
  // unchanged field (a solve per frequency or per Newton sweep on one geometry): nothing to do
  if (long(epsinv_host.size()) == n9 && std::memcmp(epsinv_host.data(), epsinv, sizeof(double) * n9) == 0) return;
  double amax = 0;
  for (long i = 0; i < n9; ++i) {
    if (!std::isfinite(epsinv[i])) throw Error(OMX_ERR_INVALID_ARG, "epsinv has non-finite entries");
    amax = std::max(amax, std::fabs(epsinv[i]));
  }
  for (long p = 0; p < N; ++p) {
    const double* e = epsinv + 9 * p;
    for (int a = 0; a < 3; ++a)
      for (int b = a + 1; b < 3; ++b)
        if (std::fabs(e[a + 3 * b] - e[b + 3 * a]) > 1e-12 * amax)
          throw Error(OMX_ERR_UNSUPPORTED,
                      "epsinv must be symmetric per pixel (the reference's eid! and _eid_g! agree only then)");
    if (!(e[0] + e[4] + e[8] > 0)) throw Error(OMX_ERR_INVALID_ARG, "epsinv trace must be positive");
  }
  epsinv_host.assign(epsinv, epsinv + n9);
  OMX_HIP_CHECK(hipSetDevice(device));
  upload_eps<float>(*this, p32);
  upload_eps<double>(*this, p64);
}

void Solver::set_option(const std::string& key, double v) {
  if (key == "fp32_tol")
    opt.fp32_tol_rel = v;
  else if (key == "fp32_tol_omega2")
    opt.fp32_tol_omega2 = v;
  else if (key == "fp32_switch")
    opt.fp32_switch = v;
  else if (key == "seed_tol")
    opt.seed_tol = std::max(0.0, v);
  else if (key == "maxiter_fp64")
    opt.maxiter_fp64 = int(v);
  else if (key == "guard")
    opt.guard = int(v);
  else if (key == "seed")
    opt.seed = uint64_t(v);
  else if (key == "plain")
    opt.plain = v != 0 ? 1 : 0;
  else if (key == "warm_start")
    opt.warm_start = v != 0 ? 1 : 0;
  else if (key == "band_guess")
    opt.band_guess = v != 0 ? 1 : 0;
  else if (key == "guess_lowres")
    opt.guess_lowres = std::max(0, int(v));
  else if (key == "refine")
    opt.refine = v != 0 ? 1 : 0;
  else if (key == "fp32_newton")
    opt.fp32_newton = v != 0 ? 1 : 0;
  else if (key == "jd_rtol")
    opt.jd_rtol = v;
  else if (key == "jd_maxit")
    opt.jd_maxit = int(v);
  else if (key == "jd_fp32")
    opt.jd_fp32 = v != 0 ? 1 : 0;
  else if (key == "jd_deflate")
    opt.jd_deflate = v != 0 ? 1 : 0;
  else if (key == "jd_first_tol")
    opt.jd_first_tol = std::max(0.0, v);
  else if (key == "streams2")
    opt.streams2 = v != 0 ? 1 : 0;
  else if (key == "graphs")
    opt.graphs = v != 0 ? 1 : 0;
  else if (key == "jac_stop32")
    opt.jac_stop32 = v;
  else if (key == "gram_blocks")
    opt.gram_blocks = std::max(0, int(v));
  else if (key == "combine_blocks")
    opt.combine_blocks = std::max(0, int(v));
  else if (key == "fft_variant32" || key == "tile_rows32" || key == "tile_cols32") {
    int& o = key == "fft_variant32" ? opt.fft_variant32 : (key == "tile_rows32" ? opt.tile_rows32 : opt.tile_cols32);
    const int nv = key == "fft_variant32" ? (v != 0 ? 1 : 0) : std::max(0, int(v));
    if (o == nv) return;
    const int old = o;
    o = nv;
    try {
      build_plan<float>(*this, p32);
    } catch (...) {
      o = old;
      build_plan<float>(*this, p32);
      throw;
    }
    workspaces.clear();  // batches size their dot partials by the operator in use
  } else if (key == "reg_ops" || key == "slab3d") {
    int& o = key == "reg_ops" ? opt.reg_ops : opt.slab3d;
    if (o == (v != 0 ? 1 : 0)) return;  // unchanged: keep the cached workspaces
    o = v != 0 ? 1 : 0;
    select_operator(*this, p32);
    select_operator(*this, p64);
    workspaces.clear();  // batches size their dot partials by the operator in use
  } else if (key == "reg_rows32" || key == "reg_rows64" || key == "reg_nt2") {
    // register-path block shapes: the default or half of it
    const int nv = int(v);
    const int full = key == "reg_rows32" ? 16 : (key == "reg_rows64" ? 8 : REG_COL_THREADS);
    if (nv != 0 && nv != full && nv != full / 2)
      throw Error(OMX_ERR_INVALID_ARG, "option '" + key + "' must be 0, " + std::to_string(full) + " or " +
                                           std::to_string(full / 2));
    int& o = key == "reg_rows32" ? opt.reg_rows32 : (key == "reg_rows64" ? opt.reg_rows64 : opt.reg_nt2);
    if (o == nv) return;
    o = nv;
    select_operator(*this, p32);
    select_operator(*this, p64);
    workspaces.clear();  // batches size their dot partials by the operator in use
  } else if (key == "fp64_only")
    opt.fp64_only = int(v);
  else if (key == "profile") {
    prof.reset();
    prof.enabled = v != 0;
  }
  else
    throw Error(OMX_ERR_INVALID_ARG, "unknown option '" + key + "'");
}

// Autotuner (the counterpart of FFTW's PATIENT planning at ModeSolver construction,
// maxwell.jl:940-943): timed trials, on this handle's grid, of the kernels an nev-mode
// solve_k spends its FP32 phase in, with the shapes that phase launches them with: nev
// band instances of nev columns, the operator on half of each block, the Hermitian Gram of
// [X R P] against [AX AR AP] and the residual block update.  Trials:
//   * operator (2D cells): the register-FFT path where the grid admits it, and the LDS
//     Stockham path for each FFT radix variant and tile shape (rows per block of passes 1/3,
//     columns per block of pass 2) that fits; a Stockham candidate is kept only if its
//     output equals the default plan's to 1e-5; 3D cells: the yz-slab and five-pass forms;
//   * the Gram row partition (gram_blocks) and the block-update partition (combine_blocks).
// On the register path, its block shapes (rows per block of passes 1/3, threads of pass 2)
// are trials too.  The fastest of each becomes the handle's option.  choice_out (9 ints):
// reg_ops, fft_variant32, tile_rows32, tile_cols32, gram_blocks, combine_blocks, slab3d,
// reg_rows32, reg_nt2; us_out (3):
// microseconds per launch of the chosen operator application, Gram and update.
void Solver::autotune(int nev, int* choice_out, double* us_out) {
  if (nev < 1 || nev > 16) throw Error(OMX_ERR_UNSUPPORTED, "nev must be in 1..16");
  if (3 * nev > K) throw Error(OMX_ERR_BAD_SHAPE, "problem too small for the requested block size");
  OMX_HIP_CHECK(hipSetDevice(device));
  const Options saved = opt;
  // on any error the handle keeps its options and plans as they were
  struct Restore {
    Solver& S;
    const Options& o;
    bool armed = true;
    ~Restore() {
      if (!armed) return;
      S.opt = o;
      try {
        build_plan<float>(S, S.p32);
        select_operator(S, S.p64);
      } catch (...) {
      }
    }
  } restore{*this, saved};
  const int M = nev, ninst = nev;
  Batch<float> b(*this, ninst, M, 2);  // not from the cache: freed when tuning ends
  hipStream_t s = stream;
  for (int q = 0; q < 2; ++q) {
    launch_randn<float>(b.p(b.X[q]), ld, b.istride, ninst, M, K, 101 + q, s);
    launch_randn<float>(b.p(b.AX[q]), ld, b.istride, ninst, M, K, 201 + q, s);
    launch_randn<float>(b.p(b.R[q]), ld, b.istride, ninst, M, K, 301 + q, s);
    launch_randn<float>(b.p(b.P[q]), ld, b.istride, ninst, M, K, 401 + q, s);
    launch_randn<float>(b.p(b.AP[q]), ld, b.istride, ninst, M, K, 501 + q, s);
  }
  launch_randn<float>(b.p(b.AR), ld, b.istride, ninst, M, K, 601, s);
  // every instance in the middle of an iteration with all M columns active
  std::vector<int> shp(size_t(F_NFIELDS) * ninst, 0);
  for (int i = 0; i < ninst; ++i) {
    shp[F_NACT * ninst + i] = M;
    shp[F_NP * ninst + i] = M;
    shp[F_NNEW * ninst + i] = M;
    shp[F_M * ninst + i] = M;
    shp[F_ONE * ninst + i] = 1;
    shp[F_NPN * ninst + i] = M;
  }
  OMX_HIP_CHECK(hipMemcpyAsync(b.shape.as<void>(), shp.data(), sizeof(int) * shp.size(), hipMemcpyHostToDevice, s));
  std::vector<double> coef(size_t(2) * ninst * b.pm * b.pm), lam(size_t(ninst) * b.pm), kh(ninst);
  uint64_t z = 0x1234567;
  for (auto& c : coef) {
    z = z * 6364136223846793005ull + 1442695040888963407ull;
    c = (double(z >> 11) * (1.0 / 9007199254740992.0) - 0.5) * 0.2;
  }
  for (size_t j = 0; j < lam.size(); ++j) lam[j] = 0.4 + 1e-3 * double(j % 16);
  for (int i = 0; i < ninst; ++i) kh[i] = 1.2 + 0.01 * i;
  OMX_HIP_CHECK(hipMemcpyAsync(b.Cx.as<void>(), coef.data(), sizeof(double) * coef.size(), hipMemcpyHostToDevice, s));
  OMX_HIP_CHECK(hipMemcpyAsync(b.lam.as<void>(), lam.data(), sizeof(double) * lam.size(), hipMemcpyHostToDevice, s));
  OMX_HIP_CHECK(hipMemcpyAsync(b.kd.as<void>(), kh.data(), sizeof(double) * ninst, hipMemcpyHostToDevice, s));
  // operator vectors: the leading half of every instance's block
  const int nop = std::max(1, M / 2);
  std::vector<int> c0(ninst, 0), nc(ninst, nop);
  OMX_HIP_CHECK(hipMemcpyAsync(b.selc0.as<void>(), c0.data(), sizeof(int) * ninst, hipMemcpyHostToDevice, s));
  OMX_HIP_CHECK(hipMemcpyAsync(b.selnc.as<void>(), nc.data(), sizeof(int) * ninst, hipMemcpyHostToDevice, s));
  OMX_HIP_CHECK(hipStreamSynchronize(s));
  hipEvent_t ea, eb;
  OMX_HIP_CHECK(hipEventCreate(&ea));
  OMX_HIP_CHECK(hipEventCreate(&eb));
  auto time_us = [&](const auto& launch) {
    launch();  // warm-up
    const int reps = 5;
    OMX_HIP_CHECK(hipEventRecord(ea, s));
    for (int r = 0; r < reps; ++r) launch();
    OMX_HIP_CHECK(hipEventRecord(eb, s));
    OMX_HIP_CHECK(hipEventSynchronize(eb));
    float ms = 0;
    OMX_HIP_CHECK(hipEventElapsedTime(&ms, ea, eb));
    return 1e3 * double(ms) / reps;
  };
  auto op_launch = [&] { b.apply(OP_M, b.X[0], b.AX[0], b.selc0.as<int>(), b.selnc.as<int>()); };
  // instance 0's operator output on the host, to check a candidate plan against the default
  const size_t nchk = size_t(nop) * ld;
  auto op_out0 = [&] {
    std::vector<c32> h(nchk);
    op_launch();
    OMX_HIP_CHECK(hipMemcpyAsync(h.data(), b.p(b.AX[0]), sizeof(c32) * nchk, hipMemcpyDeviceToHost, s));
    OMX_HIP_CHECK(hipStreamSynchronize(s));
    return h;
  };
  auto rel_diff = [&](const std::vector<c32>& a, const std::vector<c32>& r) {
    double d = 0, n = 0;
    for (int c = 0; c < nop; ++c)
      for (long e = 0; e < K; ++e) {
        const c32 x = a[size_t(c) * ld + e], y = r[size_t(c) * ld + e];
        d += double(x.x - y.x) * (x.x - y.x) + double(x.y - y.y) * (x.y - y.y);
        n += double(y.x) * y.x + double(y.y) * y.y;
      }
    return std::sqrt(d / std::max(n, 1e-300));
  };
  struct OpChoice {
    int reg, variant, rows, cols, slab, rrows = 0, nt2 = 0;
  };
  OpChoice best{saved.reg_ops, saved.fft_variant32, saved.tile_rows32, saved.tile_cols32, saved.slab3d,
                saved.reg_rows32, saved.reg_nt2};
  double best_op = 1e300;
  auto try_plan = [&](const OpChoice& c, const std::vector<c32>* ref) {
    opt.reg_ops = c.reg;
    opt.slab3d = c.slab;
    opt.fft_variant32 = c.variant;
    opt.tile_rows32 = c.rows;
    opt.tile_cols32 = c.cols;
    opt.reg_rows32 = c.rrows;
    opt.reg_nt2 = c.nt2;
    try {
      build_plan<float>(*this, p32);
    } catch (const Error&) {
      return;  // the tile or radix variant does not fit this grid's FFT bound
    }
    if (ref && rel_diff(op_out0(), *ref) > 1e-5) return;
    const double t = time_us(op_launch);
    if (t < best_op) {
      best_op = t;
      best = c;
    }
  };
  opt.reg_ops = 0;
  opt.fft_variant32 = opt.tile_rows32 = opt.tile_cols32 = 0;
  build_plan<float>(*this, p32);
  const std::vector<c32> ref = op_out0();  // default Stockham plan (or the 3D default form)
  if (nz == 1) {
    if (p32.reg_ok)
      for (int rr : {0, 16, 8})  // 0: by launch size
        for (int nt2 : {0, REG_COL_THREADS, REG_COL_THREADS / 2})
          try_plan(OpChoice{1, 0, 0, 0, saved.slab3d, rr, nt2}, &ref);
    const int budget = 3072;
    for (int variant : {0, 1})
      for (int rows : {8, 4, 2, 1}) {
        if (ny % rows != 0 || 3 * rows * nx > budget) continue;
        for (int cols : {16, 8, 4, 2, 1}) {
          if (nx % cols != 0 || 3 * cols * ny > budget) continue;
          try_plan(OpChoice{0, variant, rows, cols, saved.slab3d}, &ref);
        }
      }
  } else {
    if (p32.slab_ok) try_plan(OpChoice{saved.reg_ops, 0, 0, 0, 1}, &ref);
    try_plan(OpChoice{saved.reg_ops, 0, 0, 0, 0}, &ref);
  }
  if (best_op >= 1e300) throw Error(OMX_ERR_INTERNAL, "autotune: no operator plan passed its check");
  opt.reg_ops = best.reg;
  opt.slab3d = best.slab;
  opt.fft_variant32 = best.variant;
  opt.tile_rows32 = best.rows;
  opt.tile_cols32 = best.cols;
  opt.reg_rows32 = best.rrows;
  opt.reg_nt2 = best.nt2;
  build_plan<float>(*this, p32);
  select_operator(*this, p64);
  // partitions, with the chosen operator plan in place
  ColList<float> Y = Batch<float>::list({b.part(b.X[0], b.f(F_XOFF), b.f(F_NACT)), b.part(b.R[0], b.f(F_ROFF), b.f(F_NACT)),
                                         b.part(b.P[0], nullptr, b.f(F_NP))});
  ColList<float> AY = Batch<float>::list({b.part(b.AX[0], b.f(F_XOFF), b.f(F_NACT)),
                                          b.part(b.AR, b.f(F_ROFF), b.f(F_NACT)), b.part(b.AP[0], nullptr, b.f(F_NP))});
  int best_g = 0, best_c = 0;
  double tg = 1e300, tc = 1e300;
  const int gdef = nz > 1 ? 512 : 1024, cdef = nz > 1 ? 1024 : 2048;
  for (int gb : {128, 256, 512, 1024, 2048, 4096}) {
    opt.gram_blocks = gb;
    b.set_active(ninst);
    const double t = time_us([&] { b.gram(Y, AY, 3 * M, 3 * M, -1, false, true); });
    if (t < tg * 0.98 || (gb == gdef && t < tg * 1.02)) {  // the default wins near-ties
      tg = t;
      best_g = gb;
    }
  }
  for (int cb : {256, 512, 1024, 2048, 4096, 8192}) {
    opt.combine_blocks = cb;
    b.set_active(ninst);
    const double t = time_us([&] {
      launch_combine<float>(CMB_RESID, Y, AY, b.coef(b.Cx, b.f(F_NACT), b.lam.as<double>()), b.qmax,
                            b.out(b.X[1], b.f(F_XOFF)), b.out(b.AX[1], b.f(F_XOFF)), b.out(b.R[1], nullptr, 0), ninst,
                            K, b.nblk, b.normp.as<double>(), s);
    });
    if (t < tc * 0.98 || (cb == cdef && t < tc * 1.02)) {
      tc = t;
      best_c = cb;
    }
  }
  (void)hipEventDestroy(ea);
  (void)hipEventDestroy(eb);
  opt.gram_blocks = best_g;
  opt.combine_blocks = best_c;
  workspaces.clear();  // batches size their dot partials by the operator in use
  restore.armed = false;
  if (choice_out) {
    const int c[9] = {opt.reg_ops, opt.fft_variant32, opt.tile_rows32, opt.tile_cols32, opt.gram_blocks,
                      opt.combine_blocks, opt.slab3d, opt.reg_rows32, opt.reg_nt2};
    std::memcpy(choice_out, c, sizeof(c));
  }
  if (us_out) {
    us_out[0] = best_op;
    us_out[1] = tg;
    us_out[2] = tc;
  }
}

// Host FP64 blocks (m vectors of K complex, contiguous) -> operator -> host.
void Solver::apply(OpKind op, int prec, double k, const double* Hin, double* Hout, int m, double* dots) {
  if (m <= 0) throw Error(OMX_ERR_INVALID_ARG, "m must be positive");
  const bool dot = op == OP_MKH || op == OP_MKHW;
  if (!Hin || (!dot && !Hout) || (dot && !dots)) throw Error(OMX_ERR_INVALID_ARG, "NULL buffer");
  if (!std::isfinite(k)) throw Error(OMX_ERR_INVALID_ARG, "k must be finite");
  OMX_HIP_CHECK(hipSetDevice(device));
  const size_t nbytes64 = sizeof(c64) * size_t(m) * K;
  scratch_a.ensure(nbytes64);
  scratch_b.ensure(nbytes64);
  DevMem kbuf(sizeof(double));
  OMX_HIP_CHECK(hipMemcpyAsync(kbuf.as<void>(), &k, sizeof(double), hipMemcpyHostToDevice, stream));
  OMX_HIP_CHECK(hipMemcpyAsync(scratch_a.as<void>(), Hin, nbytes64, hipMemcpyHostToDevice, stream));
  const int dps32 = p32.plan.dot_blocks, dps64 = p64.plan.dot_blocks;
  const int dps = prec == OMX_PREC_FP32 ? dps32 : dps64;
  DevMem dotbuf(sizeof(double) * size_t(m) * dps);
  VecSel sel{1, m, nullptr, nullptr};
  if (prec == OMX_PREC_FP64) {
    DevMem Dm(sizeof(c64) * size_t(m) * 3 * N);
    BView<double> X{scratch_a.as<c64>(), K, 0}, Y{scratch_b.as<c64>(), K, 0};
    launch_operator<double>(op, p64.plan, X, Y, Dm.as<c64>(), sel, kbuf.as<double>(), dotbuf.as<double>(), dps, stream);
  } else if (prec == OMX_PREC_FP32) {
    DevMem x32(sizeof(c32) * size_t(m) * K), y32(sizeof(c32) * size_t(m) * K), Dm(sizeof(c32) * size_t(m) * 3 * N);
    launch_convert<double, float>(scratch_a.as<c64>(), K, 0, x32.as<c32>(), K, 0, 1, m, K, stream);
    BView<float> X{x32.as<c32>(), K, 0}, Y{y32.as<c32>(), K, 0};
    launch_operator<float>(op, p32.plan, X, Y, Dm.as<c32>(), sel, kbuf.as<double>(), dotbuf.as<double>(), dps, stream);
    if (!dot) launch_convert<float, double>(y32.as<c32>(), K, 0, scratch_b.as<c64>(), K, 0, 1, m, K, stream);
    OMX_HIP_CHECK(hipStreamSynchronize(stream));
  } else {
    throw Error(OMX_ERR_INVALID_ARG, "precision must be OMX_PREC_FP32 or OMX_PREC_FP64");
  }
  if (dot) {
    std::vector<double> h(size_t(m) * dps);
    OMX_HIP_CHECK(hipMemcpyAsync(h.data(), dotbuf.as<void>(), sizeof(double) * h.size(), hipMemcpyDeviceToHost, stream));
    OMX_HIP_CHECK(hipStreamSynchronize(stream));
    for (int j = 0; j < m; ++j) {
      double s = 0;
      for (int b = 0; b < dps; ++b) s += h[size_t(j) * dps + b];
      dots[j] = s;
    }
  } else {
    OMX_HIP_CHECK(hipMemcpyAsync(Hout, scratch_b.as<void>(), nbytes64, hipMemcpyDeviceToHost, stream));
    OMX_HIP_CHECK(hipStreamSynchronize(stream));
  }
}

// HMH (maxwell.jl:288-291): out[j] = Re<H_j, kx_ct(ifft(A . fft(kx_tc H_j)))>, i.e. the FP64
// operator with the per-pixel field A in place of eps^-1 (A = NULL: eps^-1 itself, so that
// out = <H, M H>), dotted with its input.
void Solver::hmh(double k, const double* A, const double* H, int m, double* out) {
  if (m <= 0) throw Error(OMX_ERR_INVALID_ARG, "m must be positive");
  if (!H || !out) throw Error(OMX_ERR_INVALID_ARG, "NULL buffer");
  if (!std::isfinite(k)) throw Error(OMX_ERR_INVALID_ARG, "k must be finite");
  OMX_HIP_CHECK(hipSetDevice(device));
  OpPlan<double> plan = p64.plan;
  DevMem a6, a6c;
  if (A) {
    check_symmetric(N, A, "A");
    upload_tensor6<double>(*this, A, a6, a6c);
    plan.eps6 = a6.as<double>();
    plan.eps6c = a6c.as<double>();
  }
  const size_t nbytes = sizeof(c64) * size_t(m) * K;
  scratch_a.ensure(nbytes);
  scratch_b.ensure(nbytes);
  DevMem kbuf(sizeof(double)), Dm(sizeof(c64) * size_t(m) * 3 * N);
  OMX_HIP_CHECK(hipMemcpyAsync(kbuf.as<void>(), &k, sizeof(double), hipMemcpyHostToDevice, stream));
  OMX_HIP_CHECK(hipMemcpyAsync(scratch_a.as<void>(), H, nbytes, hipMemcpyHostToDevice, stream));
  BView<double> X{scratch_a.as<c64>(), K, 0}, Y{scratch_b.as<c64>(), K, 0};
  launch_operator<double>(OP_M, plan, X, Y, Dm.as<c64>(), VecSel{1, m, nullptr, nullptr}, kbuf.as<double>(), nullptr,
                          0, stream);
  std::vector<double> y(size_t(2) * m * K);
  OMX_HIP_CHECK(hipMemcpyAsync(y.data(), scratch_b.as<void>(), nbytes, hipMemcpyDeviceToHost, stream));
  OMX_HIP_CHECK(hipStreamSynchronize(stream));
  for (int j = 0; j < m; ++j) {
    const double* h = H + 2 * size_t(j) * K;
    const double* v = y.data() + 2 * size_t(j) * K;
    double acc = 0;
    for (long i = 0; i < 2 * K; ++i) acc += h[i] * v[i];
    out[j] = acc;
  }
}

// group_index (lib/ModeAnalysis/src/group_index.jl:21-29) per mode j at its own k_j:
// n_g = (omega + HMH(ev_j, eps^-1 deps/domega eps^-1) / 2) / HMkH(ev_j); deps_domega NULL
// means a non-dispersive medium (the HMH term vanishes).
void Solver::group_index(double omega, int nev, const double* kmags, const double* evecs, const double* deps_domega,
                         double* ng_out) {
  if (nev <= 0 || !kmags || !evecs || !ng_out) throw Error(OMX_ERR_INVALID_ARG, "bad group_index arguments");
  if (!(omega > 0) || !std::isfinite(omega)) throw Error(OMX_ERR_INVALID_ARG, "omega must be positive and finite");
  std::vector<double> A;
  if (deps_domega) {
    check_symmetric(N, deps_domega, "deps_domega");
    A.assign(size_t(9) * N, 0.0);
    pixel_product3(N, epsinv_host.data(), deps_domega, epsinv_host.data(), A.data());
  }
  for (int j = 0; j < nev; ++j) {
    const double* ev = evecs + 2 * size_t(j) * K;
    double hk = 0, hh = 0;
    apply(OP_MKH, OMX_PREC_FP64, kmags[j], ev, nullptr, 1, &hk);
    if (deps_domega) hmh(kmags[j], A.data(), ev, 1, &hh);
    ng_out[j] = (omega + 0.5 * hh) / hk;
  }
}

namespace {

// A batch of the requested shape from the handle's workspace cache (created on first use;
// at most four shapes are kept, least recently used freed first).  Returned with cur = 0.
// The caller holds the returned pointer for as long as it uses the batch: a later request
// in the same solve may drop the entry from the cache, and the batch then lives until the
// caller lets it go.
template <typename T> std::shared_ptr<Batch<T>> workspace(Solver& S, int ninst, int M, int maxiter, int tag = 0) {
  const int tb = int(sizeof(T));
  for (size_t j = 0; j < S.workspaces.size(); ++j) {
    const Solver::Workspace& w = S.workspaces[j];
    if (w.tbytes == tb && w.ninst == ninst && w.M == M && w.maxiter >= maxiter && w.tag == tag) {
      Solver::Workspace hit = w;
      S.workspaces.erase(S.workspaces.begin() + long(j));
      S.workspaces.push_back(hit);
      std::shared_ptr<Batch<T>> b(hit.batch, static_cast<Batch<T>*>(hit.batch.get()));
      b->cur = 0;
      return b;
    }
  }
  if (S.workspaces.size() >= 4) S.workspaces.erase(S.workspaces.begin());
  std::shared_ptr<void> p(new Batch<T>(S, ninst, M, maxiter, tag == 1 ? S.stream2 : nullptr),
                          [](void* q) { delete static_cast<Batch<T>*>(q); });
  const size_t bytes = size_t(2 * tb) * size_t(ninst) * M * (11 * size_t(S.ld) + 3 * size_t(S.N));
  S.workspaces.push_back(Solver::Workspace{tb, ninst, M, maxiter, tag, p, bytes});
  return std::shared_ptr<Batch<T>>(p, static_cast<Batch<T>*>(p.get()));
}

// dst column j (stride ld) <- the low plane waves of src vector j: entry c + 2(gx + hw gy) of
// a vector of 2 hw^2 goes to c + 2(ix + nx iy) with ix = gx (gx < hw/2) or nx - hw + gx, and
// iy likewise; the other entries of the columns are left as they are
template <typename T>
__global__ __launch_bounds__(256) void k_scatter_lowres(const c64* src, int ncols, int hw, int nx, int ny,
                                                        cplx<T>* dst, long ld) {
  const long per = 2L * hw * hw, n = per * ncols;
  for (long e = long(blockIdx.x) * 256 + threadIdx.x; e < n; e += long(gridDim.x) * 256) {
    const int col = int(e / per);
    const long r = e - long(col) * per;
    const int c = int(r & 1);
    const long q = r >> 1;
    const int gx = int(q % hw), gy = int(q / hw);
    const int ix = gx < hw / 2 ? gx : nx - hw + gx;
    const int iy = gy < hw / 2 ? gy : ny - hw + gy;
    const c64 v = src[e];
    dst[long(col) * ld + 2 * (long(ix) + long(nx) * iy) + c] = cplx<T>{T(v.x), T(v.y)};
  }
}

template <typename T>
void upload_block(Solver& S, Batch<T>& b, const double* H, int ncols, int inst_from, int inst_to) {
  // H: ncols host vectors (K complex FP64 each) -> columns 0..ncols-1 of instances [from, to)
  DevMem tmp(sizeof(c64) * size_t(ncols) * S.K);
  OMX_HIP_CHECK(hipMemcpyAsync(tmp.as<void>(), H, sizeof(c64) * size_t(ncols) * S.K, hipMemcpyHostToDevice, S.stream));
  for (int i = inst_from; i < inst_to; ++i)
    launch_convert<double, T>(tmp.as<c64>(), S.K, 0, b.p(b.X[b.cur]) + i * b.istride, b.ld, 0, 1, ncols, S.K,
                              S.stream);
  OMX_HIP_CHECK(hipStreamSynchronize(S.stream));
}

// as upload_block for start vectors given by their hw x hw lowest plane waves (2 hw^2
// complex each, entry c + 2(gx + hw gy), FFT order in gx and gy); the rest of each column is 0
template <typename T>
void upload_block_lowres(Solver& S, Batch<T>& b, const double* H, int ncols, int hw, int inst_from, int inst_to) {
  const long per = 2L * hw * hw;
  DevMem tmp;
  const c64* src = S.guess_dev;
  if (!src) {
    tmp.alloc(sizeof(c64) * size_t(ncols) * per);
    OMX_HIP_CHECK(hipMemcpyAsync(tmp.as<void>(), H, sizeof(c64) * size_t(ncols) * per, hipMemcpyHostToDevice, S.stream));
    src = tmp.as<c64>();
  }
  for (int i = inst_from; i < inst_to; ++i) {
    cplx<T>* dst = b.p(b.X[b.cur]) + i * b.istride;
    OMX_HIP_CHECK(hipMemsetAsync(dst, 0, sizeof(cplx<T>) * size_t(ncols) * b.ld, S.stream));
    const long n = per * ncols;
    hipLaunchKernelGGL((k_scatter_lowres<T>), dim3(unsigned(std::min<long>((n + 255) / 256, 4096))), dim3(256), 0,
                       S.stream, src, ncols, hw, S.nx, S.ny, dst, b.ld);
  }
  OMX_HIP_CHECK(hipGetLastError());
  OMX_HIP_CHECK(hipStreamSynchronize(S.stream));
}

}  // namespace

void Solver::solve_omega2(double k, int nev, double tol, int maxiter, double* H_io, double* w2_out, omx_info* info) {
  if (nev < 1 || nev > 16) throw Error(OMX_ERR_UNSUPPORTED, "nev must be in 1..16");
  if (3 * nev > K) throw Error(OMX_ERR_BAD_SHAPE, "problem too small for the requested block size");
  if (!w2_out) throw Error(OMX_ERR_INVALID_ARG, "omega2_out is NULL");
  if (!(tol > 0) || maxiter < 1 || !std::isfinite(k)) throw Error(OMX_ERR_INVALID_ARG, "bad tol/maxiter/k");
  OMX_HIP_CHECK(hipSetDevice(device));
  hipEvent_t e0, e1;
  OMX_HIP_CHECK(hipEventCreate(&e0));
  OMX_HIP_CHECK(hipEventCreate(&e1));
  OMX_HIP_CHECK(hipEventRecord(e0, stream));
  SolveStats stats;
  std::vector<int> act{1}, bands{0}, nchk{nev};
  const auto b64p = workspace<double>(*this, 1, nev, std::max(maxiter, opt.maxiter_fp64));
  Batch<double>& b64 = *b64p;
  OMX_HIP_CHECK(hipMemcpyAsync(b64.kd.as<void>(), &k, sizeof(double), hipMemcpyHostToDevice, stream));
  if (!opt.fp64_only) {
    const auto b32p = workspace<float>(*this, 1, nev, maxiter);
    Batch<float>& b32 = *b32p;
    OMX_HIP_CHECK(hipMemcpyAsync(b32.kd.as<void>(), &k, sizeof(double), hipMemcpyHostToDevice, stream));
    if (H_io)
      upload_block<float>(*this, b32, H_io, nev, 0, 1);
    else
      launch_randn<float>(b32.p(b32.X[0]), ld, b32.istride, 1, nev, K, opt.seed, stream);
    // FP32 phase: locking threshold at the FP32 accuracy floor (relative to the Ritz values)
    const double tol32 = std::max(tol, opt.fp32_tol_omega2 * std::max(1e-300, std::fabs(k * k)));
    stats.it32 += b32.run(act, bands, nchk, {nev}, {tol32}, maxiter, stats);
    launch_convert<float, double>(b32.p(b32.X[b32.cur]), ld, b32.istride, b64.p(b64.X[0]), ld, b64.istride, 1, nev, K,
                                  stream);
  } else {
    if (H_io)
      upload_block<double>(*this, b64, H_io, nev, 0, 1);
    else
      launch_randn<double>(b64.p(b64.X[0]), ld, b64.istride, 1, nev, K, opt.seed, stream);
  }
  stats.it64 += b64.run(act, bands, nchk, {nev}, {tol}, std::max(maxiter, opt.maxiter_fp64), stats);
  const int status = b64.hst[0].status;
  auto lamv = b64.rayleigh();
  for (int j = 0; j < nev; ++j) w2_out[j] = lamv[j];
  last_evecs.ensure(sizeof(c64) * size_t(nev) * ld);
  OMX_HIP_CHECK(hipMemcpyAsync(last_evecs.as<void>(), b64.p(b64.X[b64.cur]), sizeof(c64) * size_t(nev) * ld,
                               hipMemcpyDeviceToDevice, stream));
  last_nev = nev;
  OMX_HIP_CHECK(hipEventRecord(e1, stream));
  if (H_io) {
    for (int j = 0; j < nev; ++j)
      OMX_HIP_CHECK(hipMemcpyAsync(H_io + 2 * size_t(j) * K, last_evecs.as<c64>() + size_t(j) * ld, sizeof(c64) * K,
                                   hipMemcpyDeviceToHost, stream));
  }
  OMX_HIP_CHECK(hipStreamSynchronize(stream));
  float ms = 0;
  OMX_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (info) {
    info->converged = status == 0 ? 1 : 0;
    info->n_newton = 0;
    info->lobpcg_iters_fp32 = stats.it32;
    info->lobpcg_iters_fp64 = stats.it64;
    info->n_matvec = stats.matvec;
    info->n_precon = stats.precon;
    info->max_residual = 0;
    info->device_ms = ms;
  }
  if (status == OMX_ERR_NOT_CONVERGED) last_error = "LOBPCG reached maxiter before converging";
}

namespace {

// Newton driver over a batch of band instances (instance i follows band i).
template <typename T>
int newton_phase(Solver& S, Batch<T>& b, std::vector<NewtonState>& hns, const std::vector<int>& bands,
                 const std::vector<int>& nchk, const std::vector<int>& msize, const std::vector<double>& floors,
                 int maxiter, int max_steps, SolveStats& stats, int& newton_total, bool strict, bool inexact = true,
                 std::vector<int>* inst_iters = nullptr) {
  const int ninst = b.ninst;
  std::vector<int> act(ninst);
  std::vector<double> tols(ninst);
  int iters = 0;
  for (int step = 0; step < max_steps; ++step) {
    bool any = false;
    for (int i = 0; i < ninst; ++i) {
      act[i] = hns[i].done ? 0 : 1;
      any = any || act[i];
      // inexact Newton: the eigenvalue error ~ resid^2 / gap only has to stay well below
      // the current |Δω²|, so the locking threshold follows it down to the floor
      const double f = std::fabs(hns[i].f);
      tols[i] = (inexact && hns[i].nsolves > 0 && std::isfinite(f)) ? std::max(floors[i], 0.1 * f) : floors[i];
    }
    if (!any) break;
    iters += b.run(act, bands, nchk, msize, tols, maxiter, stats);
    if (inst_iters)
      for (int i = 0; i < ninst; ++i)
        if (act[i]) (*inst_iters)[i] += b.hst[i].niter;
    b.newton(hns);
    ++newton_total;
    if (S.trace) {
      std::fprintf(stderr, "[omx] %s newton step %d:", sizeof(T) == 4 ? "fp32" : "fp64", step);
      for (int i = 0; i < ninst; ++i)
        if (act[i])
          std::fprintf(stderr, " b%d(it %d tol %.0e f %.1e)", i, b.hst[i].niter, tols[i], hns[i].f);
      std::fprintf(stderr, "\n");
    }
    // a band may only finish on an eigenpair converged to the requested tolerance: when
    // |Δω²| met k_tol on a looser (inexact-Newton) eigensolve, re-solve at the same k
    for (int i = 0; i < ninst && strict; ++i)
      if (act[i] && hns[i].done == 1 && tols[i] > floors[i]) {
        hns[i].done = 0;
        hns[i].f = 0.0;
      }
  }
  return iters;
}

template <typename T>
int newton_phase(Solver& S, Batch<T>& b, std::vector<NewtonState>& hns, const std::vector<int>& bands,
                 const std::vector<int>& nchk, const std::vector<int>& msize, double tol_floor, int maxiter,
                 int max_steps, SolveStats& stats, int& newton_total, bool strict, bool inexact = true) {
  return newton_phase(S, b, hns, bands, nchk, msize, std::vector<double>(size_t(b.ninst), tol_floor), maxiter,
                      max_steps, stats, newton_total, strict, inexact);
}

// newton_phase over the bands split into two batches on their own streams (instances
// [0, A.ninst) in A, the rest in B): the two LOBPCG runs of a Newton step advance in
// lockstep, iteration by iteration, so that the latency-bound dense steps of one batch
// overlap the memory-bound kernels of the other.  Same updates as newton_phase.
template <typename T>
int newton_phase2(Solver& S, Batch<T>& A, Batch<T>& B, std::vector<NewtonState>& hns, const std::vector<int>& bands,
                  const std::vector<int>& nchk, const std::vector<int>& msize, const std::vector<double>& floors,
                  int maxiter, int max_steps, SolveStats& stats, int& newton_total) {
  const int na = A.ninst, nb = B.ninst;
  auto part = [](const auto& v, int off, int n) {
    return std::vector<typename std::decay<decltype(v[0])>::type>(v.begin() + off, v.begin() + off + n);
  };
  const auto bA = part(bands, 0, na), bB = part(bands, na, nb), cA = part(nchk, 0, na), cB = part(nchk, na, nb);
  const auto mA = part(msize, 0, na), mB = part(msize, na, nb);
  int iters = 0;
  for (int step = 0; step < max_steps; ++step) {
    std::vector<int> act(na + nb);
    std::vector<double> tols(na + nb);
    bool any = false;
    for (int i = 0; i < na + nb; ++i) {
      act[i] = hns[i].done ? 0 : 1;
      any = any || act[i];
      const double f = std::fabs(hns[i].f);
      tols[i] = (hns[i].nsolves > 0 && std::isfinite(f)) ? std::max(floors[i], 0.1 * f) : floors[i];
    }
    if (!any) break;
    const auto aA = part(act, 0, na), aB = part(act, na, nb);
    const bool runA = std::count(aA.begin(), aA.end(), 1) > 0, runB = std::count(aB.begin(), aB.end(), 1) > 0;
    typename Batch<T>::RunState ra, rb;
    if (runA) A.run_begin(ra, aA, bA, cA, mA, part(tols, 0, na), maxiter, stats);
    if (runB) B.run_begin(rb, aB, bB, cB, mB, part(tols, na, nb), maxiter, stats);
    while ((runA && !ra.done) || (runB && !rb.done)) {
      if (runA && !ra.done) A.run_step(ra, stats);
      if (runB && !rb.done) B.run_step(rb, stats);
    }
    if (runA) iters += A.run_end(ra);
    if (runB) iters += B.run_end(rb);
    std::vector<NewtonState> hA = part(hns, 0, na), hB = part(hns, na, nb);
    if (runA) A.newton(hA);
    if (runB) B.newton(hB);
    std::copy(hA.begin(), hA.end(), hns.begin());
    std::copy(hB.begin(), hB.end(), hns.begin() + na);
    ++newton_total;
    if (S.trace) {
      std::fprintf(stderr, "[omx] fp32 newton step %d (two streams):", step);
      for (int i = 0; i < na + nb; ++i)
        if (act[i])
          std::fprintf(stderr, " b%d(it %d tol %.0e f %.1e)", i, (i < na ? A.hst[i] : B.hst[i - na]).niter, tols[i],
                       hns[i].f);
      std::fprintf(stderr, "\n");
    }
  }
  return iters;
}

using cd = std::complex<double>;

// Complex products spelled out: std::complex's operator* keeps C's infinity and NaN rules
// through a library call per product, and the values here are finite.
inline cd cmulf(cd a, cd b) {
  return cd(a.real() * b.real() - a.imag() * b.imag(), a.real() * b.imag() + a.imag() * b.real());
}
inline cd cjmulf(cd a, cd b) {  // conj(a) b
  return cd(a.real() * b.real() + a.imag() * b.imag(), a.real() * b.imag() - a.imag() * b.real());
}

// Hermitian eigenproblem by cyclic complex Jacobi (n <= RF_MAXC): A (row-major n x n,
// Hermitian) is destroyed; w ascending, V's columns the eigenvectors (row-major,
// V[r * n + c]).  The rotation of pair p < q is U = D R with D = diag(1, conj(e)) on (p, q),
// e = A[p][q] / |A[p][q]|, and R the real rotation [[c, s], [-s, c]] that zeroes the real
// 2 x 2 block [[A_pp, |A_pq|], [|A_pq|, A_qq]] (t = tan of the angle, the smaller root of
// t^2 + 2 theta t - 1 = 0 with theta = (A_qq - A_pp) / (2 |A_pq|)); A <- U^H A U, V <- V U.
// A pair is left alone once |A_pq| <= 1e-15 sqrt(|A_pp A_qq|) or <= 1e-17 |A|_F (rounding
// keeps a rotated matrix's off-diagonal entries at about that level), and the sweeps end
// with the first one that rotates no pair.
void jacobi_herm(int n, cd* A, double* w, cd* V) {
  for (int i = 0; i < n * n; ++i) V[i] = 0.0;
  for (int i = 0; i < n; ++i) V[i * n + i] = 1.0;
  double fro = 0;
  for (int i = 0; i < n * n; ++i) fro += std::norm(A[i]);
  const double floor2 = 1e-34 * fro;  // (1e-17 |A|_F)^2
  for (int sweep = 0; sweep < 40; ++sweep) {
    int rotations = 0;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        const cd apq = A[p * n + q];
        const double a2 = std::norm(apq);
        const double app = A[p * n + p].real(), aqq = A[q * n + q].real();
        if (!(a2 > floor2 && a2 > 1e-30 * std::fabs(app * aqq))) continue;
        ++rotations;
        const double a = std::sqrt(a2);
        const cd e = apq / a, ec = std::conj(e);
        const double theta = 0.5 * (aqq - app) / a;
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), sn = t * c;
        const cd sec = sn * ec, cec = c * ec, se = sn * e, ce = c * e;
        for (int k = 0; k < n; ++k) {  // columns: A <- A U
          const cd akp = A[k * n + p], akq = A[k * n + q];
          A[k * n + p] = c * akp - cmulf(sec, akq);
          A[k * n + q] = sn * akp + cmulf(cec, akq);
        }
        for (int k = 0; k < n; ++k) {  // rows: A <- U^H A
          const cd apk = A[p * n + k], aqk = A[q * n + k];
          A[p * n + k] = c * apk - cmulf(se, aqk);
          A[q * n + k] = sn * apk + cmulf(ce, aqk);
        }
        A[p * n + p] = app - t * a;
        A[q * n + q] = aqq + t * a;
        A[p * n + q] = A[q * n + p] = 0.0;
        for (int k = 0; k < n; ++k) {
          const cd vkp = V[k * n + p], vkq = V[k * n + q];
          V[k * n + p] = c * vkp - cmulf(sec, vkq);
          V[k * n + q] = sn * vkp + cmulf(cec, vkq);
        }
      }
    if (rotations == 0) break;
  }
  int idx[RF_MAXC];
  for (int i = 0; i < n; ++i) idx[i] = i;
  std::sort(idx, idx + n, [&](int a, int b) { return A[a * n + a].real() < A[b * n + b].real(); });
  cd V2[RF_MAXC * RF_MAXC];
  for (int c = 0; c < n; ++c) {
    w[c] = A[idx[c] * n + idx[c]].real();
    for (int r = 0; r < n; ++r) V2[r * n + c] = V[r * n + idx[c]];
  }
  for (int i = 0; i < n * n; ++i) V[i] = V2[i];
}

// Generalised Hermitian eigenproblem G c = lam B c for n <= RF_MAXC (G, B column-major
// with leading dimension RF_MAXC, their Hermitian parts used): B = L L^H, H = L^-1 G L^-H,
// H y = lam y by complex Jacobi, c = L^-H y.  lam ascending, C column-major n x n.  False
// when B is not numerically positive definite.
bool dense_geig(int n, const c64* Gp, const c64* Bp, std::vector<double>& lam, std::vector<cd>& C) {
  if (n < 1 || n > RF_MAXC) return false;
  auto herm = [](const c64* M, int a, int b) {
    const c64 u = M[a + RF_MAXC * b], v = M[b + RF_MAXC * a];
    return cd(0.5 * (u.x + v.x), 0.5 * (u.y - v.y));
  };
  cd L[RF_MAXC * RF_MAXC], Wm[RF_MAXC * RF_MAXC], Hm[RF_MAXC * RF_MAXC], V[RF_MAXC * RF_MAXC];
  double ld[RF_MAXC];  // the real diagonal of L
  for (int j = 0; j < n; ++j) {  // L column-major lower: B = L L^H
    double d = herm(Bp, j, j).real();
    for (int k = 0; k < j; ++k) d -= std::norm(L[j + n * k]);
    if (!(d > 1e-14)) return false;
    ld[j] = std::sqrt(d);
    L[j + n * j] = ld[j];
    for (int i = j + 1; i < n; ++i) {
      cd v = herm(Bp, i, j);
      for (int k = 0; k < j; ++k) v -= cmulf(L[i + n * k], std::conj(L[j + n * k]));
      L[i + n * j] = v / ld[j];
    }
  }
  auto fsolve = [&](cd* x) {  // x <- L^-1 x
    for (int i = 0; i < n; ++i) {
      cd v = x[i];
      for (int k = 0; k < i; ++k) v -= cmulf(L[i + n * k], x[k]);
      x[i] = v / ld[i];
    }
  };
  // W = L^-1 G (columns), H = L^-1 W^H, held row-major as Jacobi's input (H = H^H)
  cd col[RF_MAXC];
  for (int j = 0; j < n; ++j) {
    for (int i = 0; i < n; ++i) col[i] = herm(Gp, i, j);
    fsolve(col);
    for (int i = 0; i < n; ++i) Wm[i + n * j] = col[i];
  }
  for (int j = 0; j < n; ++j) {
    for (int i = 0; i < n; ++i) col[i] = std::conj(Wm[j + n * i]);
    fsolve(col);
    for (int i = 0; i < n; ++i) Hm[i * n + j] = col[i];  // H[i][j]
  }
  for (int i = 0; i < n; ++i)
    for (int j = i; j < n; ++j) {
      const cd h = 0.5 * (Hm[i * n + j] + std::conj(Hm[j * n + i]));
      Hm[i * n + j] = h;
      Hm[j * n + i] = std::conj(h);
    }
  lam.resize(n);
  jacobi_herm(n, Hm, lam.data(), V);
  // C = L^-H Y (Y = V's columns): back substitution with L^H (upper)
  C.assign(size_t(n) * n, 0.0);
  for (int j = 0; j < n; ++j) {
    for (int i = n - 1; i >= 0; --i) {
      cd v = V[i * n + j];
      for (int k = i + 1; k < n; ++k) v -= cjmulf(L[k + n * i], C[k + size_t(n) * j]);
      C[i + size_t(n) * j] = v / ld[i];
    }
  }
  return true;
}

// FP64 Newton phase by eigenpair refinement instead of LOBPCG.  Instance i holds band
// bands[i] at its own k_i with a block of msize[i] FP64 columns (from the FP32 phase).
// Per Newton step: Rayleigh-Ritz of the block at k_i, then Jacobi-Davidson steps for the
// band's Ritz pair (rho, x): MINRES on (M - rho) t = r in the complement of x (relative
// tolerance jd_rtol, at most jd_maxit iterations), t orthogonalised against the block and Rayleigh-Ritz on [block, t],
// until |M x - rho x| <= tol; then the Newton update of k from rho and HMkH(x).  The
// band's Ritz value is the bands[i]-th of the block, as LOBPCG's locking order gives it.
int refine_phase(Solver& S, Batch<double>& b, std::vector<NewtonState>& hns, const std::vector<int>& bands,
                 const std::vector<int>& msize, const std::vector<double>& tols, int max_steps, SolveStats& stats,
                 int& newton_total, double& max_resid, std::vector<double>* resid_inst = nullptr) {
  const int ninst = b.ninst;
  const long K = S.K, ld = S.ld, istride = b.istride;
  hipStream_t s = S.stream;
  const auto t_start = std::chrono::steady_clock::now();
  int jd_total = 0, minres_total = 0;
  // host wall per part (OMX_TRACE): Rayleigh-Ritz, residuals, correction solves, the rest
  double t_rr = 0, t_res = 0, t_mr = 0;
  auto tnow = [] { return std::chrono::steady_clock::now(); };
  auto tms = [](std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
  };
  const int nblk_g = int(std::max<long>(1, std::min<long>((K + 31) / 32, std::max(16, 1024 / ninst))));
  const int nblk_r = int(std::min<long>(256, (K + 255) / 256));
  RefineWorkspace& W = S.refine_ws;
  if (W.ninst < ninst || W.ld != ld || W.nblk_g < nblk_g || W.nblk_r < nblk_r || W.dps < b.dps) {
    const int n = std::max(ninst, W.ld == ld ? W.ninst : 0);
    const int ng = std::max(nblk_g, W.nblk_g), nr = std::max(nblk_r, W.nblk_r), nd = std::max(b.dps, W.dps);
    W.n1d.alloc(sizeof(int) * n);
    W.zero.alloc(sizeof(int) * n);
    W.cold.alloc(sizeof(int) * n);
    W.lamd.alloc(sizeof(double) * n);
    for (DevMem* v : {&W.Tm, &W.T2, &W.ATm, &W.XB, &W.RBm}) v->alloc(sizeof(c64) * size_t(n) * ld);
    for (DevMem* v : {&W.XBf, &W.RBf, &W.Tf}) v->alloc(sizeof(cplx<float>) * size_t(n) * ld);
    W.part.alloc(2 * sizeof(c64) * size_t(n) * ng * RF_PAIRS);  // U^H V and U^H U partials
    for (DevMem* v : {&W.Gd, &W.Bd, &W.Cd}) v->alloc(sizeof(c64) * size_t(n) * RF_PAIRS);
    W.normp.alloc(sizeof(double) * size_t(n) * nr);
    W.scl.alloc(sizeof(double) * n);
    W.dots.alloc(sizeof(double) * size_t(n) * nd);
    W.ninst = n;
    W.ld = ld;
    W.nblk_g = ng;
    W.nblk_r = nr;
    W.dps = nd;
  }
  DevMem &n1d = W.n1d, &zero = W.zero, &cold = W.cold, &lamd = W.lamd, &Tm = W.Tm, &T2 = W.T2, &ATm = W.ATm,
         &XB = W.XB, &RBm = W.RBm, &part = W.part, &Gd = W.Gd, &Bd = W.Bd, &Cd = W.Cd, &normp = W.normp,
         &scl = W.scl, &dots = W.dots;
  OMX_HIP_CHECK(hipMemsetAsync(zero.as<void>(), 0, sizeof(int) * ninst, s));
  OMX_HIP_CHECK(hipMemsetAsync(Tm.as<void>(), 0, sizeof(c64) * ninst * ld, s));
  OMX_HIP_CHECK(hipMemsetAsync(T2.as<void>(), 0, sizeof(c64) * ninst * ld, s));
  OMX_HIP_CHECK(hipMemsetAsync(ATm.as<void>(), 0, sizeof(c64) * ninst * ld, s));
  OMX_HIP_CHECK(hipMemsetAsync(XB.as<void>(), 0, sizeof(c64) * ninst * ld, s));
  OMX_HIP_CHECK(hipMemsetAsync(RBm.as<void>(), 0, sizeof(c64) * ninst * ld, s));
  OMX_HIP_CHECK(hipMemcpyAsync(cold.as<void>(), bands.data(), sizeof(int) * ninst, hipMemcpyHostToDevice, s));
  const OpPlan<double>& plan = S.p64.plan;
  std::vector<int> n1h(ninst, 0), jdh(ninst, 0);
  std::vector<double> lam_band(ninst, 0.0), rn(ninst, 0.0);
  std::vector<c64> hG(size_t(ninst) * RF_PAIRS), hB(size_t(ninst) * RF_PAIRS), hC(size_t(ninst) * RF_PAIRS);
  auto up_n1 = [&](const std::vector<int>& v) {
    OMX_HIP_CHECK(hipMemcpyAsync(n1d.as<void>(), v.data(), sizeof(int) * ninst, hipMemcpyHostToDevice, s));
  };
  // Rayleigh-Ritz of [X | t] (t only where with_t[i]) for the instances with n1h[i] > 0;
  // X, AX <- [X | t] C, [AX | At] C with C the n1h[i] lowest Ritz vectors
  auto rr = [&](const std::vector<int>& with_t) {
    const auto t0 = tnow();
    struct Acc {
      double& t;
      std::chrono::steady_clock::time_point a;
      ~Acc() { t += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - a).count(); }
    } acc{t_rr, t0};
    bool any_t = false;
    for (int i = 0; i < ninst; ++i) any_t = any_t || (n1h[i] > 0 && with_t[i]);
    const RCols Ux{b.p(b.X[b.cur]), ld, istride, any_t ? Tm.as<c64>() : nullptr};
    const RCols Ua{b.p(b.AX[b.cur]), ld, istride, any_t ? ATm.as<c64>() : nullptr};
    // Rayleigh-Ritz pair: U^H (AU) and U^H U in one pass over the rows
    launch_bgram(Ux, Ua, 0, n1d.as<int>(), ninst, K, nblk_g, part.as<c64>(), Gd.as<c64>(), s,
                 part.as<c64>() + size_t(W.ninst) * W.nblk_g * RF_PAIRS, Bd.as<c64>());
    OMX_HIP_CHECK(hipMemcpyAsync(hG.data(), Gd.as<void>(), sizeof(c64) * hG.size(), hipMemcpyDeviceToHost, s));
    OMX_HIP_CHECK(hipMemcpyAsync(hB.data(), Bd.as<void>(), sizeof(c64) * hB.size(), hipMemcpyDeviceToHost, s));
    OMX_HIP_CHECK(hipStreamSynchronize(s));
    std::fill(hC.begin(), hC.end(), c64{0, 0});
    std::vector<double> lam;
    std::vector<cd> C;
    for (int i = 0; i < ninst; ++i) {
      const int m = n1h[i];
      if (m == 0) continue;
      c64* Ci = hC.data() + size_t(i) * RF_PAIRS;
      const int n = m + ((any_t && with_t[i]) ? 1 : 0);
      if (!dense_geig(n, hG.data() + size_t(i) * RF_PAIRS, hB.data() + size_t(i) * RF_PAIRS, lam, C))
        throw Error(OMX_ERR_CHOLESKY, "refinement: the Rayleigh-Ritz basis lost rank");
      for (int j = 0; j < m; ++j)
        for (int l = 0; l < n; ++l) Ci[l + RF_MAXC * j] = c64{C[l + size_t(n) * j].real(), C[l + size_t(n) * j].imag()};
      lam_band[i] = lam[bands[i]];
    }
    OMX_HIP_CHECK(hipMemcpyAsync(Cd.as<void>(), hC.data(), sizeof(c64) * hC.size(), hipMemcpyHostToDevice, s));
    // in place: each row's inputs are read before its outputs are written (one thread per
    // row); instances outside this Rayleigh-Ritz are left as they are
    launch_bcombine(Ux, Cd.as<c64>(), n1d.as<int>(), 0, b.p(b.X[b.cur]), istride, ninst, K, s);
    launch_bcombine(Ua, Cd.as<c64>(), n1d.as<int>(), 0, b.p(b.AX[b.cur]), istride, ninst, K, s);
  };
  // band column's x (XB), residual (RBm) and residual norms for the instances in n1h
  auto band_resid = [&]() {
    const auto t0 = tnow();
    OMX_HIP_CHECK(hipMemcpyAsync(lamd.as<void>(), lam_band.data(), sizeof(double) * ninst, hipMemcpyHostToDevice, s));
    launch_band_resid(b.p(b.X[b.cur]), b.p(b.AX[b.cur]), ld, istride, cold.as<int>(), lamd.as<double>(),
                      n1d.as<int>(), ninst, K, nblk_r, XB.as<c64>(), RBm.as<c64>(), normp.as<double>(), s);
    std::vector<double> h(size_t(ninst) * nblk_r);
    OMX_HIP_CHECK(hipMemcpyAsync(h.data(), normp.as<void>(), sizeof(double) * h.size(), hipMemcpyDeviceToHost, s));
    OMX_HIP_CHECK(hipStreamSynchronize(s));
    for (int i = 0; i < ninst; ++i) rn[i] = n1h[i] ? std::sqrt(sum_part(h.data() + size_t(i) * nblk_r, nblk_r)) : 0.0;
    t_res += tms(t0, tnow());
  };
  const int jd_max = 8;
  int outer = 0;
  // each band's residual at its last Newton update (the first step's is only the looser
  // jd_first_tol one); their maximum is the solve's final residual
  std::vector<double> last_rn(ninst, 0.0);
  for (int step = 0; step < max_steps; ++step) {
    bool any = false;
    for (int i = 0; i < ninst; ++i) {
      n1h[i] = hns[i].done ? 0 : msize[i];
      any = any || n1h[i] > 0;
    }
    if (!any) break;
    up_n1(n1h);
    // AX = M X at the current k, then Rayleigh-Ritz of the block
    b.apply(OP_M, b.X[b.cur], b.AX[b.cur], zero.as<int>(), n1d.as<int>());
    for (int i = 0; i < ninst; ++i) stats.matvec += n1h[i];
    rr(std::vector<int>(ninst, 0));
    ++outer;
    std::vector<int> conv(ninst, 0);
    // the first step corrects only until rho is good enough for a Newton update of k (the
    // residual at the new k is then of the size of the k change and is corrected next step)
    std::vector<double> step_tol(tols);
    if (step == 0 && S.opt.jd_first_tol > 0)
      for (int i = 0; i < ninst; ++i) step_tol[i] = std::max(tols[i], S.opt.jd_first_tol * hns[i].omega2);
    for (int jd = 0;; ++jd) {
      band_resid();
      bool need = false;
      for (int i = 0; i < ninst; ++i) {
        conv[i] = n1h[i] > 0 && rn[i] <= step_tol[i];
        jdh[i] = (n1h[i] > 0 && !conv[i]) ? 1 : 0;
        need = need || jdh[i];
      }
      if (S.trace) {
        std::fprintf(stderr, "[omx] fp64 refine step %d, correction %d, residuals:", step, jd);
        for (int i = 0; i < ninst; ++i)
          if (n1h[i]) std::fprintf(stderr, " %.1e", rn[i]);
        std::fprintf(stderr, "\n");
      }
      if (!need || jd >= jd_max) break;
      // t: (M - rho) t = r on the complement of x (the sign of t is immaterial)
      VecOps V(S, s, ninst, ld, K, XB.as<c64>());
      int its;
      const auto t_mr0 = tnow();
      if (S.opt.jd_fp32) {
        // the correction only needs jd_rtol (5e-3) relative accuracy: FP32 vectors and
        // operator, FP64 recurrences; the residual that decides convergence stays FP64
        using cf = cplx<float>;
        launch_convert<double, float>(XB.as<c64>(), ld, ld, W.XBf.as<cf>(), ld, ld, ninst, 1, K, s);
        launch_convert<double, float>(RBm.as<c64>(), ld, ld, W.RBf.as<cf>(), ld, ld, ninst, 1, K, s);
        VecOpsT<float> V32(S, s, ninst, ld, K, W.XBf.as<cf>());
        if (S.opt.jd_deflate) {
          // the block after its Rayleigh-Ritz step: orthonormal Ritz vectors, the band's x
          // among them, and r orthogonal to all of them
          const long qis = long(b.M) * ld;
          W.Xf.ensure(sizeof(cf) * size_t(ninst) * qis);
          W.nqd.ensure(sizeof(int) * size_t(ninst));
          launch_convert<double, float>(b.p(b.X[b.cur]), ld, istride, W.Xf.as<cf>(), ld, qis, ninst, b.M, K, s);
          OMX_HIP_CHECK(hipMemcpyAsync(W.nqd.as<void>(), msize.data(), sizeof(int) * ninst, hipMemcpyHostToDevice, s));
          V32.set_block(W.Xf.as<cf>(), qis, W.nqd.as<int>(), *std::max_element(msize.begin(), msize.end()));
        }
        its = minres_complement_t<float>(S, V32, lam_band, b.kd.as<double>(), W.RBf.as<cf>(), W.Tf.as<cf>(), jdh,
                                         S.opt.jd_rtol, S.opt.jd_maxit, false);
        launch_convert<float, double>(W.Tf.as<cf>(), ld, ld, Tm.as<c64>(), ld, ld, ninst, 1, K, s);
      } else {
        its = minres_complement(S, V, lam_band, b.kd.as<double>(), RBm.as<c64>(), Tm.as<c64>(), jdh, S.opt.jd_rtol,
                                S.opt.jd_maxit, false);
      }
      for (int i = 0; i < ninst; ++i) {
        stats.matvec += long(its) * jdh[i];
        stats.precon += long(its) * jdh[i];
      }
      ++jd_total;
      minres_total += its;
      OMX_HIP_CHECK(hipStreamSynchronize(s));
      t_mr += tms(t_mr0, tnow());
      // t <- t - X (X^H t), then unit norm (instances with jdh only).  Twice, unless the
      // correction solve already ran on the complement of the FP32 copy of the whole block:
      // t is then orthogonal to X up to FP32 rounding, and one FP64 pass removes that.
      std::vector<int> n1j(ninst);
      for (int i = 0; i < ninst; ++i) n1j[i] = jdh[i] ? msize[i] : 0;
      up_n1(n1j);
      const int orth_passes = (S.opt.jd_fp32 && S.opt.jd_deflate) ? 1 : 2;
      for (int pass = 0; pass < orth_passes; ++pass) {
        const RCols Xc{b.p(b.X[b.cur]), ld, istride, nullptr}, Tc{nullptr, ld, ld, Tm.as<c64>()};
        launch_bgram(Xc, Tc, 1, n1d.as<int>(), ninst, K, nblk_g, part.as<c64>(), Gd.as<c64>(), s);
        OMX_HIP_CHECK(hipMemcpyAsync(hG.data(), Gd.as<void>(), sizeof(c64) * hG.size(), hipMemcpyDeviceToHost, s));
        OMX_HIP_CHECK(hipStreamSynchronize(s));
        std::fill(hC.begin(), hC.end(), c64{0, 0});
        for (int i = 0; i < ninst; ++i) {
          if (!jdh[i]) continue;
          c64* Ci = hC.data() + size_t(i) * RF_PAIRS;
          for (int l = 0; l < msize[i]; ++l) {
            const c64 g = hG[size_t(i) * RF_PAIRS + l];  // (l, 0): <x_l, t>
            Ci[l] = c64{-g.x, -g.y};
          }
          Ci[msize[i]] = c64{1, 0};
        }
        OMX_HIP_CHECK(hipMemcpyAsync(Cd.as<void>(), hC.data(), sizeof(c64) * hC.size(), hipMemcpyHostToDevice, s));
        const RCols XT{b.p(b.X[b.cur]), ld, istride, Tm.as<c64>()};
        launch_bcombine(XT, Cd.as<c64>(), n1d.as<int>(), 1, T2.as<c64>(), ld, ninst, K, s);
        std::swap(Tm, T2);
      }
      {
        V.vdot(Tm.as<c64>(), Tm.as<c64>(), n1d.as<int>() /* nonzero = active */);
        std::vector<double> h(size_t(ninst) * V.nblk), sc(ninst, 1.0);
        OMX_HIP_CHECK(hipMemcpyAsync(h.data(), V.part.as<void>(), sizeof(double) * h.size(), hipMemcpyDeviceToHost, s));
        OMX_HIP_CHECK(hipStreamSynchronize(s));
        for (int i = 0; i < ninst; ++i)
          if (jdh[i]) {
            const double nt = std::sqrt(sum_part(h.data() + size_t(i) * V.nblk, V.nblk));
            sc[i] = nt > 0 ? 1.0 / nt : 0.0;
          }
        OMX_HIP_CHECK(hipMemcpyAsync(scl.as<void>(), sc.data(), sizeof(double) * ninst, hipMemcpyHostToDevice, s));
        launch_scale_cols(Tm.as<c64>(), ld, scl.as<double>(), n1d.as<int>(), ninst, K, s);
      }
      // At = M t
      {
        const VecSel sel{ninst, 1, zero.as<int>(), n1d.as<int>()};
        BView<double> Tv{Tm.as<c64>(), ld, ld}, ATv{ATm.as<c64>(), ld, ld};
        launch_operator<double>(OP_M, plan, Tv, ATv, b.p(b.D), sel, b.kd.as<double>(), nullptr, 0, s);
        for (int i = 0; i < ninst; ++i) stats.matvec += jdh[i];
      }
      // Rayleigh-Ritz on [X | t] for the instances still refining
      n1h = n1j;
      rr(jdh);
      ++outer;
    }
    // restore the step's instance set (those converged in the JD loop were dropped from n1h)
    for (int i = 0; i < ninst; ++i) n1h[i] = hns[i].done ? 0 : msize[i];
    up_n1(n1h);
    band_resid();
    // Newton update from rho and HMkH(x) of the band vector
    {
      const VecSel sel{ninst, 1, zero.as<int>(), n1d.as<int>()};
      BView<double> XBv{XB.as<c64>(), ld, ld};
      launch_operator<double>(OP_MKH, plan, XBv, XBv, b.p(b.D), sel, b.kd.as<double>(), dots.as<double>(), b.dps, s);
      std::vector<double> h(size_t(ninst) * b.dps);
      OMX_HIP_CHECK(hipMemcpyAsync(h.data(), dots.as<void>(), sizeof(double) * h.size(), hipMemcpyDeviceToHost, s));
      OMX_HIP_CHECK(hipStreamSynchronize(s));
      std::vector<double> kh(ninst);
      for (int i = 0; i < ninst; ++i) {
        NewtonState& ns = hns[i];
        kh[i] = ns.k;
        if (!n1h[i]) continue;
        const double slope = 2.0 * sum_part(h.data() + size_t(i) * b.dps, b.dps);
        ns.lam = lam_band[i];
        ns.f = lam_band[i] - ns.omega2;
        ns.slope = slope;
        ns.nsolves += 1;
        last_rn[i] = rn[i];
        if (resid_inst) (*resid_inst)[i] = rn[i];
        if (rn[i] <= tols[i] && std::fabs(ns.f) <= ns.k_tol) {
          ns.done = 1;
        } else if (ns.nsolves >= ns.max_solves) {
          ns.done = 2;
        } else {
          const double kn = ns.k - ns.f / slope;
          ns.k = (std::isfinite(kn) && kn > 0) ? kn : ns.k;
        }
        kh[i] = ns.k;
      }
      OMX_HIP_CHECK(hipMemcpyAsync(b.kd.as<void>(), kh.data(), sizeof(double) * ninst, hipMemcpyHostToDevice, s));
    }
    ++newton_total;
    if (S.trace) {
      std::fprintf(stderr, "[omx] fp64 refine step %d:", step);
      for (int i = 0; i < ninst; ++i)
        if (n1h[i]) std::fprintf(stderr, " b%d(res %.1e f %.1e)", i, rn[i], hns[i].f);
      std::fprintf(stderr, "\n");
    }
  }
  for (int i = 0; i < ninst; ++i) max_resid = std::max(max_resid, last_rn[i]);
  if (S.trace)
    std::fprintf(stderr, "[omx] fp64 refine: %.2f ms host wall (Rayleigh-Ritz %.2f, residuals %.2f, correction solves "
                 "%.2f), %d Rayleigh-Ritz steps, %d correction solves, %d MINRES iterations\n",
                 tms(t_start, tnow()), t_rr, t_res, t_mr, outer, jd_total, minres_total);
  return outer;
}

}  // namespace

void Solver::solve_k(double omega, int nev, double k_tol, double eig_tol, int max_eigsolves, int maxiter,
                     const double* kguess, const double* Hguess, int nguess, double* kmags, double* evecs,
                     omx_info* info) {
  if (nev < 1 || nev > 16) throw Error(OMX_ERR_UNSUPPORTED, "nev must be in 1..16");
  if (3 * nev > K) throw Error(OMX_ERR_BAD_SHAPE, "problem too small for the requested block size");
  if (!kmags) throw Error(OMX_ERR_INVALID_ARG, "kmags_out is NULL");
  if (!(omega > 0) || !std::isfinite(omega) || !(k_tol > 0) || !(eig_tol > 0) || max_eigsolves < 1 || maxiter < 1)
    throw Error(OMX_ERR_INVALID_ARG, "bad omega/k_tol/eig_tol/max_eigsolves/maxiter");
  if (Hguess && nguess < 0) throw Error(OMX_ERR_INVALID_ARG, "nguess must be >= 0");
  OMX_HIP_CHECK(hipSetDevice(device));
  hipEvent_t e0, e1;
  OMX_HIP_CHECK(hipEventCreate(&e0));
  OMX_HIP_CHECK(hipEventCreate(&e1));
  OMX_HIP_CHECK(hipEventRecord(e0, stream));
  const auto t_start = std::chrono::steady_clock::now();
  auto t_fp32_end = t_start;
  SolveStats stats;
  const int ninst = nev, M = nev;
  // k_guess (MaxwellEigenmodes.jl:63): omega * sqrt(1 / min_a min eps^-1[a,a,:,:])
  double k0;
  if (kguess) {
    k0 = kguess[0];
    if (!std::isfinite(k0)) throw Error(OMX_ERR_INVALID_ARG, "kguess must be finite");
  } else {
    double mn = 1e300;
    for (long p = 0; p < N; ++p)
      for (int a = 0; a < 3; ++a) mn = std::min(mn, epsinv_host[9 * p + 4 * a]);
    k0 = omega * std::sqrt(1.0 / mn);
  }
  const double w2 = omega * omega;
  std::vector<NewtonState> hns(ninst);
  for (int i = 0; i < ninst; ++i) {
    NewtonState& s = hns[i];
    std::memset(&s, 0, sizeof(s));
    s.k = k0;
    s.omega2 = w2;
    s.band = i;
    s.max_solves = max_eigsolves;
  }
  // band i is instance i: it needs the i+1 lowest pairs locked (DFTK n_conv_check); its
  // block carries `guard` further columns (all nev columns when guard < 0)
  std::vector<int> bands(ninst), nchk(ninst), msize(ninst);
  for (int i = 0; i < ninst; ++i) {
    bands[i] = i;
    nchk[i] = i + 1;
    msize[i] = (opt.guard < 0 || opt.plain) ? M : std::min(M, i + 1 + opt.guard);
  }
  const int maxit64 = std::max(maxiter, opt.maxiter_fp64);
  const auto b64p = workspace<double>(*this, ninst, M, maxit64);
  Batch<double>& b64 = *b64p;
  int newton_total = 0;
  const int nguess_used = (Hguess && nguess > 0) ? std::min(nguess, M) : 0;
  const int ghw = opt.guess_lowres;
  if (nguess_used && ghw > 0 && (nz > 1 || ghw % 2 != 0 || ghw > nx || ghw > ny))
    throw Error(OMX_ERR_INVALID_ARG, "guess_lowres must be even and at most Nx, Ny (2D cells)");
  // start vectors into instances [from, to): full vectors, or their lowest plane waves
  auto put_guess = [&](auto& b, int from, int to) {
    using TT = typename std::remove_reference<decltype(b)>::type::value_type;
    if (ghw > 0)
      upload_block_lowres<TT>(*this, b, Hguess, nguess_used, ghw, from, to);
    else
      upload_block<TT>(*this, b, Hguess, nguess_used, from, to);
  };
  // FP32 locking threshold: loose where the FP64 refinement finishes the vectors, at the FP32
  // accuracy floor where an FP64 LOBPCG does
  const double tol32 = std::max(eig_tol, (opt.refine ? opt.fp32_tol_rel : opt.fp32_tol_omega2) * w2);

  // first eigensolve, shared by every band: the lowest nev pairs at k0 (instance 0);
  // each band then starts Newton from its own first-order estimate of k with this block
  auto seed_bands = [&](auto& b, double tol0) {
    using TT = typename std::remove_reference<decltype(b)>::type::value_type;
    std::vector<int> act(ninst, 0), n0(ninst, M), m0(ninst, M);
    act[0] = 1;
    std::vector<double> t0(ninst, tol0);
    std::vector<double> kh(ninst, k0);
    OMX_HIP_CHECK(hipMemcpyAsync(b.kd.template as<void>(), kh.data(), sizeof(double) * ninst, hipMemcpyHostToDevice,
                                 stream));
    const int it = b.run(act, bands, n0, m0, t0, sizeof(TT) == 4 ? maxiter : maxit64, stats);
    auto lam = b.rayleigh();
    auto hk = b.hmkh_cols(0, M);
    for (int i = 0; i < ninst; ++i) {
      const double f = lam[i] - w2, slope = 2.0 * hk[i];
      hns[i].f = f;
      hns[i].lam = lam[i];
      hns[i].slope = slope;
      hns[i].nsolves = 1;
      if (std::fabs(f) <= hns[i].k_tol) {
        hns[i].done = 1;
      } else {
        const double kn = k0 - f / slope;
        hns[i].k = std::isfinite(kn) && kn > 0 ? kn : k0;
      }
      kh[i] = hns[i].k;
    }
    for (int i = 1; i < ninst; ++i)
      OMX_HIP_CHECK(hipMemcpyAsync(b.p(b.X[b.cur]) + i * b.istride, b.p(b.X[b.cur]), sizeof(cplx<TT>) * b.istride,
                                   hipMemcpyDeviceToDevice, stream));
    OMX_HIP_CHECK(hipMemcpyAsync(b.kd.template as<void>(), kh.data(), sizeof(double) * ninst, hipMemcpyHostToDevice,
                                 stream));
    ++newton_total;
    return it;
  };

  // previous-solve warm start (the EME slice chain, eme.jl:58-76): every band starts at the
  // previous solve's k of that band with the previous eigenvectors as its block (kept on
  // the device), and the shared first eigensolve is skipped
  const bool warm = opt.warm_start && !opt.plain && !kguess && !Hguess && last_nev > 0 &&
                    int(last_kmags.size()) >= last_nev;
  // per-band guesses (the learned warm start): band i starts at kguess[i] with the Hguess
  // block (filled up with random columns), and the shared first eigensolve is skipped
  const bool banded = opt.band_guess && !opt.plain && kguess && nguess_used > 0;
  if (banded)
    for (int i = 0; i < ninst; ++i)
      if (!(kguess[i] > 0) || !std::isfinite(kguess[i]))
        throw Error(OMX_ERR_INVALID_ARG, "band_guess needs a positive finite kguess per band");
  auto guided_block = [&](auto& b) {
    using TT = typename std::remove_reference<decltype(b)>::type::value_type;
    if (nguess_used < M) launch_randn<TT>(b.p(b.X[b.cur]), ld, b.istride, ninst, M, K, opt.seed, stream);
    put_guess(b, 0, ninst);
    std::vector<double> kh(ninst);
    for (int i = 0; i < ninst; ++i) {
      kh[i] = kguess[i];
      hns[i].k = kh[i];
    }
    OMX_HIP_CHECK(hipMemcpyAsync(b.kd.template as<void>(), kh.data(), sizeof(double) * ninst, hipMemcpyHostToDevice,
                                 stream));
    OMX_HIP_CHECK(hipStreamSynchronize(stream));
  };
  auto warm_block = [&](auto& b) {
    using TT = typename std::remove_reference<decltype(b)>::type::value_type;
    const int nw = std::min(last_nev, M);
    if (nw < M) launch_randn<TT>(b.p(b.X[b.cur]), ld, b.istride, ninst, M, K, opt.seed, stream);
    for (int i = 0; i < ninst; ++i)
      launch_convert<double, TT>(last_evecs.as<c64>(), ld, 0, b.p(b.X[b.cur]) + i * b.istride, b.ld, 0, 1, nw, K,
                                 stream);
    std::vector<double> kh(ninst);
    for (int i = 0; i < ninst; ++i) {
      kh[i] = last_kmags[std::min(i, nw - 1)];
      hns[i].k = kh[i];
    }
    OMX_HIP_CHECK(hipMemcpyAsync(b.kd.template as<void>(), kh.data(), sizeof(double) * ninst, hipMemcpyHostToDevice,
                                 stream));
    OMX_HIP_CHECK(hipStreamSynchronize(stream));
  };

  if (opt.plain) {
    // the reference's algorithm (solve.jl:89-113 with the DFTK backend): each band runs
    // Newton from k_guess with its own random nev-column start, every eigensolve to eig_tol
    std::vector<double> kh(ninst, k0);
    OMX_HIP_CHECK(hipMemcpyAsync(b64.kd.as<void>(), kh.data(), sizeof(double) * ninst, hipMemcpyHostToDevice, stream));
    for (auto& st : hns) st.k_tol = k_tol;
    // one band at a time (a single solve per band, as the reference runs them); the band's
    // start block goes into the batch's current buffer when its turn comes
    for (int b = 0; b < ninst; ++b) {
      launch_randn<double>(b64.p(b64.X[b64.cur]) + b * b64.istride, ld, b64.istride, 1, M, K, opt.seed + 7919u * b,
                           stream);
      if (nguess_used) put_guess(b64, b, b + 1);
      const std::vector<NewtonState> keep = hns;
      for (int i = 0; i < ninst; ++i)
        if (i != b) hns[i].done = 3;  // not this band's turn
      int steps = 0;
      stats.it64 +=
          newton_phase(*this, b64, hns, bands, nchk, msize, eig_tol, maxit64, max_eigsolves, stats, steps, false, false);
      newton_total = std::max(newton_total, steps);
      for (int i = 0; i < ninst; ++i)
        if (i != b) hns[i] = keep[i];
    }
  } else if (!opt.fp64_only) {
    // two half-batches of bands on two streams (the seed eigensolve runs in the first)
    const int na = (ninst + 1) / 2, nb = ninst - na;
    const bool split = opt.streams2 && nb > 0 && !warm && !banded && opt.fp32_newton && stream2;
    const auto b32p = workspace<float>(*this, split ? na : ninst, M, maxiter);
    Batch<float>& b32 = *b32p;
    const std::shared_ptr<Batch<float>> b32q = split ? workspace<float>(*this, nb, M, maxiter, 1) : nullptr;
    for (auto& s : hns) s.k_tol = std::max(k_tol, opt.fp32_switch * w2);
    if (warm) {
      warm_block(b32);
    } else if (banded) {
      guided_block(b32);
    } else if (split) {
      Batch<float>& B = *b32q;
      launch_randn<float>(b32.p(b32.X[0]), ld, b32.istride, 1, M, K, opt.seed, stream);
      if (nguess_used) put_guess(b32, 0, 1);
      // the shared first eigensolve in instance 0 of the first half, as seed_bands
      std::vector<int> act(na, 0), n0(na, M), m0(na, M), bA(bands.begin(), bands.begin() + na);
      act[0] = 1;
      std::vector<double> t0(na, std::max(tol32, seed_tol_rel(true) * w2)), kh(ninst, k0);
      OMX_HIP_CHECK(hipMemcpyAsync(b32.kd.as<void>(), kh.data(), sizeof(double) * na, hipMemcpyHostToDevice, stream));
      stats.it32 += b32.run(act, bA, n0, m0, t0, maxiter, stats);
      const auto lam = b32.rayleigh();
      const auto hk = b32.hmkh_cols(0, M);
      for (int i = 0; i < ninst; ++i) {
        const double f = lam[i] - w2, slope = 2.0 * hk[i];
        hns[i].f = f;
        hns[i].lam = lam[i];
        hns[i].slope = slope;
        hns[i].nsolves = 1;
        if (std::fabs(f) <= hns[i].k_tol) {
          hns[i].done = 1;
        } else {
          const double kn = k0 - f / slope;
          hns[i].k = std::isfinite(kn) && kn > 0 ? kn : k0;
        }
        kh[i] = hns[i].k;
      }
      const cplx<float>* x0 = b32.p(b32.X[b32.cur]);
      for (int i = 1; i < na; ++i)
        OMX_HIP_CHECK(hipMemcpyAsync(b32.p(b32.X[b32.cur]) + i * b32.istride, x0, sizeof(c32) * b32.istride,
                                     hipMemcpyDeviceToDevice, stream));
      B.cur = 0;
      for (int i = 0; i < nb; ++i)
        OMX_HIP_CHECK(hipMemcpyAsync(B.p(B.X[0]) + i * B.istride, x0, sizeof(c32) * B.istride, hipMemcpyDeviceToDevice,
                                     stream));
      OMX_HIP_CHECK(hipMemcpyAsync(b32.kd.as<void>(), kh.data(), sizeof(double) * na, hipMemcpyHostToDevice, stream));
      OMX_HIP_CHECK(hipMemcpyAsync(B.kd.as<void>(), kh.data() + na, sizeof(double) * nb, hipMemcpyHostToDevice, stream));
      OMX_HIP_CHECK(hipStreamSynchronize(stream));  // the second stream starts from these
      ++newton_total;
    } else {
      launch_randn<float>(b32.p(b32.X[0]), ld, b32.istride, 1, M, K, opt.seed, stream);
      if (nguess_used) put_guess(b32, 0, 1);
      stats.it32 += seed_bands(b32, std::max(tol32, seed_tol_rel() * w2));
    }
    if (split) {
      stats.it32 += newton_phase2(*this, b32, *b32q, hns, bands, nchk, msize, std::vector<double>(ninst, tol32), maxiter,
                                  max_eigsolves - 1, stats, newton_total);
      OMX_HIP_CHECK(hipStreamSynchronize(stream2));
      Batch<float>& B = *b32q;
      launch_convert<float, double>(b32.p(b32.X[b32.cur]), ld, b32.istride, b64.p(b64.X[0]), ld, b64.istride, na, M, K,
                                    stream);
      launch_convert<float, double>(B.p(B.X[B.cur]), ld, B.istride, b64.p(b64.X[0]) + na * b64.istride, ld, b64.istride,
                                    nb, M, K, stream);
      OMX_HIP_CHECK(hipMemcpyAsync(b64.kd.as<void>(), b32.kd.as<void>(), sizeof(double) * na, hipMemcpyDeviceToDevice,
                                   stream));
      OMX_HIP_CHECK(hipMemcpyAsync(b64.kd.as<double>() + na, B.kd.as<void>(), sizeof(double) * nb,
                                   hipMemcpyDeviceToDevice, stream));
    } else {
      if (opt.fp32_newton)
        stats.it32 += newton_phase(*this, b32, hns, bands, nchk, msize, tol32, maxiter, max_eigsolves - 1, stats,
                                   newton_total, false);
      launch_convert<float, double>(b32.p(b32.X[b32.cur]), ld, b32.istride, b64.p(b64.X[0]), ld, b64.istride, ninst, M,
                                    K, stream);
      OMX_HIP_CHECK(hipMemcpyAsync(b64.kd.as<void>(), b32.kd.as<void>(), sizeof(double) * ninst, hipMemcpyDeviceToDevice,
                                   stream));
    }
    OMX_HIP_CHECK(hipStreamSynchronize(stream));
    t_fp32_end = std::chrono::steady_clock::now();
    for (auto& s : hns) {
      if (s.done == 1) s.done = 0;  // refine in FP64
      s.k_tol = k_tol;
    }
  } else {
    for (auto& s : hns) s.k_tol = k_tol;
    if (warm) {
      warm_block(b64);
    } else if (banded) {
      guided_block(b64);
    } else {
      launch_randn<double>(b64.p(b64.X[0]), ld, b64.istride, 1, M, K, opt.seed, stream);
      if (nguess_used) put_guess(b64, 0, 1);
      stats.it64 += seed_bands(b64, std::max(eig_tol, seed_tol_rel() * w2));
    }
  }
  double max_resid = 0;
  if (!opt.plain) {
    const int left = std::max(1, max_eigsolves - newton_total);
    if (opt.refine && !opt.fp64_only)
      stats.it64 += refine_phase(*this, b64, hns, bands, msize, std::vector<double>(size_t(ninst), eig_tol), left,
                                 stats, newton_total, max_resid);
    else
      stats.it64 +=
          newton_phase(*this, b64, hns, bands, nchk, msize, eig_tol, maxit64, left, stats, newton_total, true);
  }
  if (trace) {
    OMX_HIP_CHECK(hipStreamSynchronize(stream));
    const auto t1 = std::chrono::steady_clock::now();
    std::fprintf(stderr, "[omx] solve_k phases: fp32 %.2f ms, fp64 %.2f ms\n",
                 std::chrono::duration<double, std::milli>(t_fp32_end - t_start).count(),
                 std::chrono::duration<double, std::milli>(t1 - t_fp32_end).count());
  }
  bool ok = true;
  for (int i = 0; i < ninst; ++i) {
    kmags[i] = hns[i].k;
    if (hns[i].done != 1) ok = false;
  }
  // eigenvector of band i = column i of instance i; canonicalise phase (fields.jl:258-260)
  last_evecs.ensure(sizeof(c64) * size_t(nev) * ld);
  for (int i = 0; i < ninst; ++i)
    OMX_HIP_CHECK(hipMemcpyAsync(last_evecs.as<c64>() + size_t(i) * ld, b64.p(b64.X[b64.cur]) + i * b64.istride + i * ld,
                                 sizeof(c64) * K, hipMemcpyDeviceToDevice, stream));
  last_nev = nev;
  last_kmags.assign(kmags, kmags + nev);
  canonicalize_phase(b64.kd.as<double>(), last_evecs.as<c64>(), ld, ld, nev, 1);
  OMX_HIP_CHECK(hipEventRecord(e1, stream));
  if (evecs) get_evecs(evecs);
  OMX_HIP_CHECK(hipStreamSynchronize(stream));
  float ms = 0;
  OMX_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (info) {
    info->converged = ok ? 1 : 0;
    info->n_newton = newton_total;
    info->lobpcg_iters_fp32 = stats.it32;
    info->lobpcg_iters_fp64 = stats.it64;
    info->n_matvec = stats.matvec;
    info->n_precon = stats.precon;
    info->max_residual = max_resid;
    info->device_ms = ms;
  }
  if (!ok) throw Error(OMX_ERR_NOT_CONVERGED, "solve_k: Newton iteration did not converge within max_eigsolves");
}

void Solver::get_evecs(double* evecs) {
  if (!evecs) throw Error(OMX_ERR_INVALID_ARG, "evecs_out is NULL");
  if (last_nev == 0) throw Error(OMX_ERR_INVALID_ARG, "no eigenvectors on this handle yet");
  for (int j = 0; j < last_nev; ++j)
    OMX_HIP_CHECK(hipMemcpyAsync(evecs + 2 * size_t(j) * K, last_evecs.as<c64>() + size_t(j) * ld, sizeof(c64) * K,
                                 hipMemcpyDeviceToHost, stream));
  OMX_HIP_CHECK(hipStreamSynchronize(stream));
}

namespace {

// Host conversion of a (3,3,Nx,Ny) column-major field to the operator's two 6-entry layouts,
// written to device memory at rowm / colm (2D cells).
template <typename T> void put_tensor6(const Solver& S, const double* E, T* rowm, T* colm, hipStream_t s) {
  const long N = S.N;
  std::vector<T> e6(size_t(6) * N), e6c(size_t(6) * N);
  for (long p = 0; p < N; ++p) {
    const double* e = E + 9 * p;
    T* o = e6.data() + 6 * p;
    o[0] = T(e[0]);
    o[1] = T(e[4]);
    o[2] = T(e[8]);
    o[3] = T(e[3]);
    o[4] = T(e[6]);
    o[5] = T(e[7]);
  }
  for (int ix = 0; ix < S.nx; ++ix)
    for (int iy = 0; iy < S.ny; ++iy)
      for (int a = 0; a < 6; ++a) e6c[6 * (iy + size_t(S.ny) * ix) + a] = e6[6 * (ix + size_t(S.nx) * iy) + a];
  OMX_HIP_CHECK(hipMemcpyAsync(rowm, e6.data(), sizeof(T) * e6.size(), hipMemcpyHostToDevice, s));
  OMX_HIP_CHECK(hipMemcpyAsync(colm, e6c.data(), sizeof(T) * e6c.size(), hipMemcpyHostToDevice, s));
  OMX_HIP_CHECK(hipStreamSynchronize(s));  // the host arrays go out of scope
}

}  // namespace

// Several solve_k problems on this grid in one batch (no reference counterpart as one call:
// the reference sweeps run solve_k, src/solve.jl:182-196, one problem per call or process,
// lib/ModeSweeps/src/batch.jl:256-265).  Problem p's band j is instance off_p + j; the block
// size is the largest nev, each instance using min(nev_p, j + 1 + guard) columns.  Every
// launch of the FP32 phase (shared first eigensolve per problem, inexact Newton), of the
// FP64 refinement and of the phase canonicalisation covers all instances; the operator
// reads instance i's eps^-1 through plan.eps_slot.  Per problem the results are those of
// omx_solve_k on that problem alone, up to rounding.
void Solver::solve_k_multi(std::vector<MultiProblem>& probs) {
  require_2d("the batched solve_k");
  const int P = int(probs.size());
  if (P == 0) return;
  int ninst = 0, M = 0, maxiter = 1, max_eigsolves = 1;
  for (auto& p : probs) {
    if (p.nev < 1 || p.nev > 16) throw Error(OMX_ERR_UNSUPPORTED, "nev must be in 1..16");
    if (3 * p.nev > K) throw Error(OMX_ERR_BAD_SHAPE, "problem too small for the requested block size");
    if (!p.kmags || !p.epsinv) throw Error(OMX_ERR_INVALID_ARG, "NULL epsinv or kmags_out");
    if (!(p.omega > 0) || !std::isfinite(p.omega) || !(p.k_tol > 0) || !(p.eig_tol > 0) || p.max_eigsolves < 1 ||
        p.maxiter < 1)
      throw Error(OMX_ERR_INVALID_ARG, "bad omega/k_tol/eig_tol/max_eigsolves/maxiter");
    check_symmetric(N, p.epsinv, "epsinv");
    ninst += p.nev;
    M = std::max(M, p.nev);
    maxiter = std::max(maxiter, p.maxiter);
    max_eigsolves = std::max(max_eigsolves, p.max_eigsolves);
  }
  OMX_HIP_CHECK(hipSetDevice(device));
  hipStream_t s = stream;
  hipEvent_t e0, e1;
  OMX_HIP_CHECK(hipEventCreate(&e0));
  OMX_HIP_CHECK(hipEventCreate(&e1));
  OMX_HIP_CHECK(hipEventRecord(e0, s));
  // the problems' eps^-1, and which one each instance reads
  const size_t n6 = size_t(6) * N;
  multi_e32r.ensure(sizeof(float) * n6 * P);
  multi_e32c.ensure(sizeof(float) * n6 * P);
  multi_e64r.ensure(sizeof(double) * n6 * P);
  multi_e64c.ensure(sizeof(double) * n6 * P);
  multi_slot.ensure(sizeof(int) * ninst);
  std::vector<int> off(P), slot(ninst);
  std::vector<double> k0(P);
  for (int p = 0, o = 0; p < P; ++p) {
    off[p] = o;
    for (int j = 0; j < probs[p].nev; ++j) slot[o + j] = p;
    o += probs[p].nev;
    put_tensor6<float>(*this, probs[p].epsinv, multi_e32r.as<float>() + n6 * p, multi_e32c.as<float>() + n6 * p, s);
    put_tensor6<double>(*this, probs[p].epsinv, multi_e64r.as<double>() + n6 * p, multi_e64c.as<double>() + n6 * p, s);
    double mn = 1e300;  // k_guess (MaxwellEigenmodes.jl:63)
    for (long q = 0; q < N; ++q)
      for (int a = 0; a < 3; ++a) mn = std::min(mn, probs[p].epsinv[9 * q + 4 * a]);
    k0[p] = probs[p].omega * std::sqrt(1.0 / mn);
  }
  OMX_HIP_CHECK(hipMemcpyAsync(multi_slot.as<void>(), slot.data(), sizeof(int) * ninst, hipMemcpyHostToDevice, s));
  OMX_HIP_CHECK(hipStreamSynchronize(s));
  // the plans read the stacked fields for the rest of this call
  struct PlanSwap {
    Solver& S;
    OpPlan<float> a;
    OpPlan<double> b;
    ~PlanSwap() {
      S.p32.plan.eps6 = a.eps6;
      S.p32.plan.eps6c = a.eps6c;
      S.p32.plan.eps_slot = nullptr;
      S.p32.plan.eps_stride = 0;
      S.p64.plan.eps6 = b.eps6;
      S.p64.plan.eps6c = b.eps6c;
      S.p64.plan.eps_slot = nullptr;
      S.p64.plan.eps_stride = 0;
    }
  } swap{*this, p32.plan, p64.plan};
  p32.plan.eps6 = multi_e32r.as<float>();
  p32.plan.eps6c = multi_e32c.as<float>();
  p32.plan.eps_slot = multi_slot.as<int>();
  p32.plan.eps_stride = long(n6);
  p64.plan.eps6 = multi_e64r.as<double>();
  p64.plan.eps6c = multi_e64c.as<double>();
  p64.plan.eps_slot = multi_slot.as<int>();
  p64.plan.eps_stride = long(n6);

  SolveStats stats;
  int newton_total = 0;
  std::vector<NewtonState> hns(ninst);
  std::vector<int> bands(ninst), nchk(ninst), msize(ninst), prob_of(ninst);
  std::vector<double> w2(ninst), tol32(ninst), tol_seed(ninst), eig_tol(ninst), kinit(ninst);
  for (int p = 0; p < P; ++p)
    for (int j = 0; j < probs[p].nev; ++j) {
      const int i = off[p] + j;
      const MultiProblem& q = probs[p];
      NewtonState& st = hns[i];
      std::memset(&st, 0, sizeof(st));
      st.k = k0[p];
      st.omega2 = q.omega * q.omega;
      st.band = j;
      st.max_solves = q.max_eigsolves;
      prob_of[i] = p;
      bands[i] = j;
      nchk[i] = j + 1;
      msize[i] = opt.guard < 0 ? q.nev : std::min(q.nev, j + 1 + opt.guard);
      w2[i] = st.omega2;
      tol32[i] = std::max(q.eig_tol, (opt.refine ? opt.fp32_tol_rel : opt.fp32_tol_omega2) * w2[i]);
      eig_tol[i] = q.eig_tol;
      kinit[i] = k0[p];
    }
  const int maxit64 = std::max(maxiter, opt.maxiter_fp64);
  const auto b64p = workspace<double>(*this, ninst, M, maxit64);
  Batch<double>& b64 = *b64p;
  std::vector<int> it32_inst(ninst, 0);

  // the shared first eigensolve of every problem (instance off_p, all nev_p pairs at k0_p);
  // each band then starts Newton from its first-order estimate with that block
  auto seed = [&](auto& b, const std::vector<double>& tfloor, int mi) {
    using TT = typename std::remove_reference<decltype(b)>::type::value_type;
    std::vector<int> act(ninst, 0), n0(ninst, M), m0(ninst, M), ncol(ninst, 0);
    std::vector<double> t0(ninst, 1.0);
    for (int p = 0; p < P; ++p) {
      const int i0 = off[p];
      act[i0] = 1;
      n0[i0] = m0[i0] = ncol[i0] = probs[p].nev;
      t0[i0] = std::max(tfloor[i0], seed_tol_rel() * w2[i0]);
    }
    OMX_HIP_CHECK(hipMemcpyAsync(b.kd.template as<void>(), kinit.data(), sizeof(double) * ninst, hipMemcpyHostToDevice, s));
    const int it = b.run(act, bands, n0, m0, t0, mi, stats);
    for (int p = 0; p < P; ++p) it32_inst[off[p]] += sizeof(TT) == 4 ? b.hst[off[p]].niter : 0;
    const auto lam = b.rayleigh();
    const auto hk = b.hmkh_multi(ncol);
    std::vector<double> kh(ninst);
    for (int p = 0; p < P; ++p) {
      const int i0 = off[p];
      for (int j = 0; j < probs[p].nev; ++j) {
        NewtonState& st = hns[i0 + j];
        const double l = lam[size_t(i0) * M + j], f = l - st.omega2, slope = 2.0 * hk[size_t(i0) * M + j];
        st.f = f;
        st.lam = l;
        st.slope = slope;
        st.nsolves = 1;
        if (std::fabs(f) <= st.k_tol) {
          st.done = 1;
        } else {
          const double kn = k0[p] - f / slope;
          st.k = std::isfinite(kn) && kn > 0 ? kn : k0[p];
        }
        kh[i0 + j] = st.k;
      }
      for (int j = 1; j < probs[p].nev; ++j)
        OMX_HIP_CHECK(hipMemcpyAsync(b.p(b.X[b.cur]) + (i0 + j) * b.istride, b.p(b.X[b.cur]) + i0 * b.istride,
                                     sizeof(cplx<TT>) * b.istride, hipMemcpyDeviceToDevice, s));
    }
    OMX_HIP_CHECK(hipMemcpyAsync(b.kd.template as<void>(), kh.data(), sizeof(double) * ninst, hipMemcpyHostToDevice, s));
    ++newton_total;
    return it;
  };

  if (!opt.fp64_only) {
    const auto b32p = workspace<float>(*this, ninst, M, maxiter);
    Batch<float>& b32 = *b32p;
    for (int i = 0; i < ninst; ++i) hns[i].k_tol = std::max(probs[prob_of[i]].k_tol, opt.fp32_switch * w2[i]);
    launch_randn<float>(b32.p(b32.X[0]), ld, b32.istride, ninst, M, K, opt.seed, s);
    stats.it32 += seed(b32, tol32, maxiter);
    if (opt.fp32_newton)
      stats.it32 += newton_phase(*this, b32, hns, bands, nchk, msize, tol32, maxiter, max_eigsolves - 1, stats,
                                 newton_total, false, true, &it32_inst);
    launch_convert<float, double>(b32.p(b32.X[b32.cur]), ld, b32.istride, b64.p(b64.X[0]), ld, b64.istride, ninst, M,
                                  K, s);
    OMX_HIP_CHECK(hipMemcpyAsync(b64.kd.as<void>(), b32.kd.as<void>(), sizeof(double) * ninst, hipMemcpyDeviceToDevice, s));
    OMX_HIP_CHECK(hipStreamSynchronize(s));
    for (int i = 0; i < ninst; ++i) {
      if (hns[i].done == 1) hns[i].done = 0;  // refine in FP64
      hns[i].k_tol = probs[prob_of[i]].k_tol;
    }
  } else {
    for (int i = 0; i < ninst; ++i) hns[i].k_tol = probs[prob_of[i]].k_tol;
    launch_randn<double>(b64.p(b64.X[0]), ld, b64.istride, ninst, M, K, opt.seed, s);
    stats.it64 += seed(b64, eig_tol, maxit64);
  }
  double max_resid = 0;
  std::vector<double> resid(ninst, 0.0);
  const int left = std::max(1, max_eigsolves - newton_total);
  if (opt.refine && !opt.fp64_only)
    stats.it64 += refine_phase(*this, b64, hns, bands, msize, eig_tol, left, stats, newton_total, max_resid, &resid);
  else
    stats.it64 += newton_phase(*this, b64, hns, bands, nchk, msize, eig_tol, maxit64, left, stats, newton_total, true);
  // band eigenvectors of every instance, phase-canonicalised together (fields.jl:258-260)
  multi_evecs.ensure(sizeof(c64) * size_t(ninst) * ld);
  for (int i = 0; i < ninst; ++i)
    OMX_HIP_CHECK(hipMemcpyAsync(multi_evecs.as<c64>() + size_t(i) * ld,
                                 b64.p(b64.X[b64.cur]) + i * b64.istride + long(bands[i]) * ld, sizeof(c64) * K,
                                 hipMemcpyDeviceToDevice, s));
  canonicalize_phase(b64.kd.as<double>(), multi_evecs.as<c64>(), ld, ld, ninst, 1);
  OMX_HIP_CHECK(hipEventRecord(e1, s));
  for (int p = 0; p < P; ++p)
    if (probs[p].evecs)
      for (int j = 0; j < probs[p].nev; ++j)
        OMX_HIP_CHECK(hipMemcpyAsync(probs[p].evecs + 2 * size_t(j) * K, multi_evecs.as<c64>() + size_t(off[p] + j) * ld,
                                     sizeof(c64) * K, hipMemcpyDeviceToHost, s));
  OMX_HIP_CHECK(hipStreamSynchronize(s));
  float ms = 0;
  OMX_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  for (int p = 0; p < P; ++p) {
    MultiProblem& q = probs[p];
    bool ok = true;
    int newton = 0, it32 = 0;
    double rmax = 0;
    for (int j = 0; j < q.nev; ++j) {
      const int i = off[p] + j;
      q.kmags[j] = hns[i].k;
      ok = ok && hns[i].done == 1;
      newton = std::max(newton, hns[i].nsolves);
      it32 += it32_inst[i];
      rmax = std::max(rmax, resid[i]);
    }
    q.info.converged = ok ? 1 : 0;
    q.info.n_newton = newton;
    q.info.lobpcg_iters_fp32 = it32;
    q.info.lobpcg_iters_fp64 = stats.it64;  // the batch's refinement steps
    q.info.n_matvec = stats.matvec;         // the batch's operator applications
    q.info.n_precon = stats.precon;
    q.info.max_residual = opt.refine && !opt.fp64_only ? rmax : 0.0;
    q.info.device_ms = ms;                  // the batch's device time
    q.status = ok ? OMX_OK : OMX_ERR_NOT_CONVERGED;
  }
}

void Solver::require_2d(const char* what) const {
  if (nz > 1) throw Error(OMX_ERR_UNSUPPORTED, std::string(what) + " is built for 2D cells only");
}

// Period sensitivity on a 3D cell (lib/MaxwellEigenmodes/src/grads/period.jl:88-90, with
// HMkH_weighted :69-74): per mode j at its own k_j, dw2dk[j] = 2 HMkH(ev_j) and
// dw2dL[j] = 2 HMkH_weighted(ev_j, g_z / Dz), both from the FP64 operator on the GPU.
void Solver::period_derivatives(int nev, const double* kmags, const double* evecs, double* dw2dk, double* dw2dL) {
  if (nz <= 1) throw Error(OMX_ERR_UNSUPPORTED, "period derivatives need a 3D cell (Nz > 1)");
  if (nev < 1 || !kmags || !evecs || !dw2dk || !dw2dL) throw Error(OMX_ERR_INVALID_ARG, "bad period_derivatives arguments");
  for (int j = 0; j < nev; ++j) {
    if (!std::isfinite(kmags[j])) throw Error(OMX_ERR_INVALID_ARG, "k must be finite");
    const double* ev = evecs + 2 * size_t(j) * K;
    double hk = 0, hl = 0;
    apply(OP_MKH, OMX_PREC_FP64, kmags[j], ev, nullptr, 1, &hk);
    apply(OP_MKHW, OMX_PREC_FP64, kmags[j], ev, nullptr, 1, &hl);
    dw2dk[j] = 2.0 * hk;
    dw2dL[j] = 2.0 * hl;
  }
}

}  // namespace omx
