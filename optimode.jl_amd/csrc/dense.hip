// Small dense FP64 kernels (see dense.hpp).  One workgroup per instance; matrices live
// in LDS, row-major, element (r, c) at M[r*ld + c].
#include "dense.hpp"
#include "omx_errors.hpp"

namespace omx {

namespace {

// One wavefront per instance: the matrices here have at most 48 x 16 entries, so the
// steps are latency-bound and a single-wave workgroup makes each barrier nearly free.
constexpr int DBLK = 64;
constexpr int DMAX = 48;       // max Y columns (3 x 16)
constexpr int DLD = DMAX + 1;  // padded LDS row stride of Y-sized matrices
constexpr int SLD = 17;        // LDS row stride of matrices with <= 16 columns

__device__ __forceinline__ c64 cadd(c64 a, c64 b) { return c64{a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ c64 csub(c64 a, c64 b) { return c64{a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ c64 cmul(c64 a, c64 b) { return c64{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ __forceinline__ c64 ccj(c64 a) { return c64{a.x, -a.y}; }
__device__ __forceinline__ c64 cscale(double s, c64 a) { return c64{s * a.x, s * a.y}; }

// deterministic pseudo-random normal (used by drop!, dftk.jl:224-234)
__device__ __forceinline__ double hash_normal(uint64_t key) {
  uint64_t z = key + 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  z ^= z >> 31;
  uint64_t y = z + 0x9e3779b97f4a7c15ull;
  y = (y ^ (y >> 30)) * 0xbf58476d1ce4e5b9ull;
  y = (y ^ (y >> 27)) * 0x94d049bb133111ebull;
  y ^= y >> 31;
  const double u1 = (double(z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  const double u2 = double(y >> 11) * (1.0 / 9007199254740992.0);
  return sqrt(-2.0 * log(u1)) * cospi(2.0 * u2);
}

// ----------------------------------------------------------------------------------------
// block-wide helpers
// ----------------------------------------------------------------------------------------

// out (ca x cb) = A^H B, A rows x ca, B rows x cb
__device__ void d_ahb(const c64* A, int lda, const c64* B, int ldb, int rows, int ca, int cb, c64* out, int ldo) {
  for (int e = threadIdx.x; e < ca * cb; e += blockDim.x) {
    const int a = e / cb, b = e - (e / cb) * cb;
    double sx = 0, sy = 0;
    for (int r = 0; r < rows; ++r) {
      const c64 x = A[r * lda + a], y = B[r * ldb + b];
      sx += x.x * y.x + x.y * y.y;
      sy += x.x * y.y - x.y * y.x;
    }
    out[a * ldo + b] = c64{sx, sy};
  }
  __syncthreads();
}

// X (rows x cx) -= A (rows x k) * B (k x cx)
__device__ void d_sub_ab(c64* X, int ldx, const c64* A, int lda, const c64* B, int ldb, int rows, int k, int cx) {
  for (int e = threadIdx.x; e < rows * cx; e += blockDim.x) {
    const int r = e / cx, c = e - (e / cx) * cx;
    c64 s{0, 0};
    for (int l = 0; l < k; ++l) s = cadd(s, cmul(A[r * lda + l], B[l * ldb + c]));
    X[r * ldx + c] = csub(X[r * ldx + c], s);
  }
  __syncthreads();
}

// X (rows x c) <- X * T (c x c), using scratch W (rows x c)
__device__ void d_rmul(c64* X, int ldx, const c64* Tm, int ldt, int rows, int c, c64* W, int ldw) {
  for (int e = threadIdx.x; e < rows * c; e += blockDim.x) {
    const int r = e / c, q = e - (e / c) * c;
    c64 s{0, 0};
    for (int l = 0; l < c; ++l) s = cadd(s, cmul(X[r * ldx + l], Tm[l * ldt + q]));
    W[r * ldw + q] = s;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < rows * c; e += blockDim.x) {
    const int r = e / c, q = e - (e / c) * c;
    X[r * ldx + q] = W[r * ldw + q];
  }
  __syncthreads();
}

__shared__ double s_red[DBLK];

__device__ double d_sum(double v) {
  s_red[threadIdx.x] = v;
  __syncthreads();
  for (int s = DBLK / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) s_red[threadIdx.x] += s_red[threadIdx.x + s];
    __syncthreads();
  }
  const double r = s_red[0];
  __syncthreads();
  return r;
}

// Upper Cholesky O = U^H U in place (O becomes U, lower part zeroed).  Returns true on
// success (all pivots positive and finite).
__device__ bool d_chol(c64* O, int ld, int n) {
  __shared__ int ok;
  if (threadIdx.x == 0) ok = 1;
  __syncthreads();
  for (int k = 0; k < n; ++k) {
    if (threadIdx.x == 0) {
      const double d = O[k * ld + k].x;
      if (!(d > 0.0) || !isfinite(d)) {
        ok = 0;
      } else {
        O[k * ld + k] = c64{sqrt(d), 0.0};
      }
    }
    __syncthreads();
    if (!ok) break;
    const double inv = 1.0 / O[k * ld + k].x;
    for (int j = k + 1 + threadIdx.x; j < n; j += blockDim.x) O[k * ld + j] = cscale(inv, O[k * ld + j]);
    __syncthreads();
    // trailing: O[i][j] -= conj(U[k][i]) U[k][j], i, j > k, upper part only
    const int m = n - k - 1;
    for (int e = threadIdx.x; e < m * m; e += blockDim.x) {
      const int i = k + 1 + e / m, j = k + 1 + (e - (e / m) * m);
      if (j >= i) O[i * ld + j] = csub(O[i * ld + j], cmul(ccj(O[k * ld + i]), O[k * ld + j]));
    }
    __syncthreads();
  }
  const int res = ok;
  for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
    const int i = e / n, j = e - (e / n) * n;
    if (j < i) O[i * ld + j] = c64{0, 0};
  }
  __syncthreads();
  return res != 0;
}

// Inverse of upper-triangular U (n x n) into W.
__device__ void d_triinv(const c64* U, int ldu, c64* W, int ldw, int n) {
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    for (int i = 0; i < n; ++i) W[i * ldw + j] = c64{0, 0};
    // solve U x = e_j, x[i] = 0 for i > j
    for (int i = j; i >= 0; --i) {
      c64 s = (i == j) ? c64{1, 0} : c64{0, 0};
      for (int l = i + 1; l <= j; ++l) s = csub(s, cmul(U[i * ldu + l], W[l * ldw + j]));
      // divide by U[i][i] (real positive)
      W[i * ldw + j] = cscale(1.0 / U[i * ldu + i].x, s);
    }
  }
  __syncthreads();
}

// normest (dftk.jl:146): max |diag| + Frobenius norm of the off-diagonal part
__device__ double d_normest(const c64* M, int ld, int n) {
  double off = 0, dmax = 0;
  for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
    const int i = e / n, j = e - (e / n) * n;
    const c64 v = M[i * ld + j];
    if (i == j)
      dmax = fmax(dmax, sqrt(v.x * v.x + v.y * v.y));
    else
      off += v.x * v.x + v.y * v.y;
  }
  // max via sum of per-thread maxima is not a max: reduce max separately
  s_red[threadIdx.x] = dmax;
  __syncthreads();
  for (int s = DBLK / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) s_red[threadIdx.x] = fmax(s_red[threadIdx.x], s_red[threadIdx.x + s]);
    __syncthreads();
  }
  const double mx = s_red[0];
  __syncthreads();
  return mx + sqrt(d_sum(off));
}

// ortho!(X; tol) (dftk.jl:151-221) on X (rows x c) in LDS.  Returns growth factor, or a
// negative value on failure.  O, W: c x c scratch.
__device__ double d_ortho(c64* X, int ldx, int rows, int c, double eps, double tol, c64* O, c64* W, c64* Wr) {
  double growth = 1.0;
  int nchol = 0;
  for (;;) {
    d_ahb(X, ldx, X, ldx, rows, c, c, O, SLD);
    double tr = 0;
    for (int i = 0; i < c; ++i) tr += O[i * SLD + i].x;
    // keep a copy for shifting
    for (int e = threadIdx.x; e < c * c; e += blockDim.x) W[(e / c) * SLD + e % c] = O[(e / c) * SLD + e % c];
    __syncthreads();
    bool success = d_chol(O, SLD, c);
    if (success) {
      ++nchol;
    } else {
      double alpha = 100.0;
      int nbad = 0;
      for (;;) {
        for (int e = threadIdx.x; e < c * c; e += blockDim.x) {
          const int i = e / c, j = e - (e / c) * c;
          O[i * SLD + j] = W[i * SLD + j];
          if (i == j) O[i * SLD + j].x += alpha * eps * tr;
        }
        __syncthreads();
        for (int e = threadIdx.x; e < c; e += blockDim.x) W[e * SLD + e].x += alpha * eps * tr;
        __syncthreads();
        alpha *= 10;
        if (d_chol(O, SLD, c)) {
          ++nchol;
          break;
        }
        if (++nbad > 10) return -1.0;
      }
    }
    d_triinv(O, SLD, W, SLD, c);
    d_rmul(X, ldx, W, SLD, rows, c, Wr, SLD);
    const double ninv = d_normest(W, SLD, c);
    growth *= ninv;
    const double condR = d_normest(O, SLD, c) * ninv;
    if (success && eps * condR * condR < tol) break;
    if (nchol > 10) return -1.0;
  }
  return growth;
}

// ortho!(X, Y, BY=Y; tol) (dftk.jl:238-285) with X rows x cx, Y rows x cy.
__device__ int d_ortho_against(c64* X, int rows, int cx, const c64* Y, int cy, double eps, double tol, c64* B,
                               c64* O, c64* W, c64* Wr, uint64_t seed) {
  // normalise columns
  for (int c = threadIdx.x; c < cx; c += blockDim.x) {
    double s = 0;
    for (int r = 0; r < rows; ++r) s += X[r * SLD + c].x * X[r * SLD + c].x + X[r * SLD + c].y * X[r * SLD + c].y;
    const double inv = 1.0 / sqrt(s);
    for (int r = 0; r < rows; ++r) X[r * SLD + c] = cscale(inv, X[r * SLD + c]);
  }
  __syncthreads();
  __shared__ int dropped[DMAX];
  __shared__ int ndropped;
  for (int niter = 1;; ++niter) {
    d_ahb(Y, SLD, X, SLD, rows, cy, cx, B, SLD);   // BYX = Y^H X
    d_sub_ab(X, SLD, Y, SLD, B, SLD, rows, cy, cx);  // X -= Y BYX
    // drop!: columns with norm <= tol are replaced by random vectors
    if (threadIdx.x == 0) ndropped = 0;
    __syncthreads();
    for (int c = threadIdx.x; c < cx; c += blockDim.x) {
      double s = 0;
      for (int r = 0; r < rows; ++r) s += X[r * SLD + c].x * X[r * SLD + c].x + X[r * SLD + c].y * X[r * SLD + c].y;
      dropped[c] = sqrt(s) <= tol ? 1 : 0;
    }
    __syncthreads();
    if (threadIdx.x == 0)
      for (int c = 0; c < cx; ++c) ndropped += dropped[c];
    __syncthreads();
    if (ndropped) {
      for (int e = threadIdx.x; e < rows * cx; e += blockDim.x) {
        const int r = e / cx, c = e - (e / cx) * cx;
        if (dropped[c]) {
          const uint64_t key = seed * 7919ull + uint64_t(niter) * 104729ull + uint64_t(r * DMAX + c) * 2;
          X[r * SLD + c] = c64{hash_normal(key), hash_normal(key + 1)};
        }
      }
      __syncthreads();
      // X[:, dropped] -= Y (Y^H X[:, dropped]) -- projecting every column again is
      // equivalent for the kept columns up to O(eps) and keeps the code uniform
      d_ahb(Y, SLD, X, SLD, rows, cy, cx, O, SLD);
      for (int e = threadIdx.x; e < rows * cx; e += blockDim.x) {
        const int r = e / cx, c = e - (e / cx) * cx;
        if (dropped[c]) {
          c64 s{0, 0};
          for (int l = 0; l < cy; ++l) s = cadd(s, cmul(Y[r * SLD + l], O[l * SLD + c]));
          X[r * SLD + c] = csub(X[r * SLD + c], s);
        }
      }
      __syncthreads();
    }
    double nb = 0;
    for (int e = threadIdx.x; e < cy * cx; e += blockDim.x) {
      const c64 v = B[(e / cx) * SLD + e % cx];
      nb += v.x * v.x + v.y * v.y;
    }
    nb = sqrt(d_sum(nb));
    if (nb < tol && niter > 1) break;
    const double growth = d_ortho(X, SLD, rows, cx, eps, tol, O, W, Wr);
    if (growth < 0) return OMX_ERR_CHOLESKY;
    if (growth * eps < tol) break;
    if (niter > 10) return OMX_ERR_CHOLESKY;
  }
  return 0;
}

// Parallel cyclic Jacobi for Hermitian A (n x n).  Round-robin pairing gives m/2
// disjoint rotations per round (m = n rounded up to even; an odd n gets a zero row and
// column, which never rotates).  The rotation of pair (p, q) zeroes A[p][q]:
//   rows  A'[p] = c A[p] - s A[q],          A'[q] = conj(s) A[p] + c A[q]
//   cols  A'[:,p] = c A[:,p] - conj(s) A[:,q],  A'[:,q] = s A[:,p] + c A[:,q]
// so A' = L A L^H with L block diagonal.  A round is two LDS passes: the m/2 rotations
// (one thread each), then the (m/2)^2 2x2 blocks of A (read from one buffer, written to
// the other) and the m x m/2 row pairs of V (in place).  A sweep is m - 1 rounds.  Sweeps
// stop after one whose rotations all have |sin| <= stop (the off-diagonal part left is then
// below about stop^2 |A|: 1e-18 for the FP64 batches' 1e-9, 1e-8 for the FP32 batches' 1e-4)
// or that performs none.  Returns the buffer holding the final
// matrix (eigenvalues on its diagonal); V holds the eigenvectors in its columns.
__device__ __forceinline__ void jac_pair(int t, int r, int m, int& p, int& q) {
  const int ia = (t == 0) ? 0 : 1 + (t - 1 + r) % (m - 1);
  const int tb = m - 1 - t;
  const int ib = 1 + (tb - 1 + r) % (m - 1);
  p = ia < ib ? ia : ib;
  q = ia < ib ? ib : ia;
}

// c = cos(theta), s = sin(theta) e^{i arg A[p][q]} with tan(2 theta) = 2|A[p][q]| / (A[q][q] - A[p][p])
// and |theta| <= pi/4: cos(2 theta) = |z| h, sin(2 theta) = h with z = (A[q][q] - A[p][p]) / (2|A[p][q]|)
// and h = 1/sqrt(1 + z^2), so c = sqrt((1 + cos 2theta)/2) >= 1/sqrt(2) and |s| = sin(2theta)/(2c),
// without cancellation.  Identity when |A[p][q]| is negligible.  Returns |s|, or -1 for
// the identity.
__device__ __forceinline__ double jac_rot(const c64* A, int p, int q, double& c, c64& s) {
  const double app = A[p * DLD + p].x, aqq = A[q * DLD + q].x;
  const c64 apq = A[p * DLD + q];
  const double a2 = apq.x * apq.x + apq.y * apq.y;
  c = 1.0;
  s = c64{0.0, 0.0};
  if (!(a2 > 1e-300 && a2 > 1e-32 * fabs(app * aqq))) return -1.0;
  const double ia = rsqrt(a2);
  const double z = 0.5 * (aqq - app) * ia;
  const double az = fabs(z);
  double sn;
  if (az > 1e150) {
    sn = 0.5 / az;
  } else {
    const double h = rsqrt(1.0 + z * z);
    const double x = 0.5 * (1.0 + az * h);
    const double rx = rsqrt(x);
    c = x * rx;
    sn = 0.5 * h * rx;
  }
  const double f = (z >= 0 ? sn : -sn) * ia;
  s = c64{f * apq.x, f * apq.y};
  return sn;
}

__device__ c64* d_jacobi(c64* A0, c64* A1, c64* V, int n, double stop) {
  __shared__ int rot[3], big[3];
  __shared__ int rp[DMAX / 2], rq[DMAX / 2];
  __shared__ double rc[DMAX / 2];
  __shared__ c64 rs[DMAX / 2];
  const int m = n + (n & 1);
  const int npair = m / 2;
  for (int e = threadIdx.x; e < m * m; e += blockDim.x) {
    const int i = e / m, j = e - (e / m) * m;
    V[i * DLD + j] = c64{i == j ? 1.0 : 0.0, 0.0};
    if (i == n || j == n) A0[i * DLD + j] = c64{0.0, 0.0};
  }
  if (threadIdx.x < 3) {
    rot[threadIdx.x] = 0;
    big[threadIdx.x] = 0;
  }
  __syncthreads();
  c64* Ac = A0;
  c64* An = A1;
  const int nblk = npair * npair, nitem = nblk + m * npair;
  for (int sweep = 0; sweep < 40; ++sweep) {
    const int sl = sweep % 3;
    if (threadIdx.x == 0) {
      rot[(sweep + 1) % 3] = 0;
      big[(sweep + 1) % 3] = 0;
    }
    for (int r = 0; r < m - 1; ++r) {
      if (threadIdx.x < npair) {
        const int t = threadIdx.x;
        int p, q;
        jac_pair(t, r, m, p, q);
        double c;
        c64 sv;
        const double sn = jac_rot(Ac, p, q, c, sv);
        if (sn >= 0.0) rot[sl] = 1;
        if (sn > stop) big[sl] = 1;
        rp[t] = p;
        rq[t] = q;
        rc[t] = c;
        rs[t] = sv;
      }
      __syncthreads();
      for (int idx = threadIdx.x; idx < nitem; idx += blockDim.x) {
        if (idx < nblk) {
          const int t1 = idx / npair, t2 = idx - (idx / npair) * npair;
          const int p1 = rp[t1], q1 = rq[t1], p2 = rp[t2], q2 = rq[t2];
          const double c1 = rc[t1], c2 = rc[t2];
          const c64 s1 = rs[t1], s2 = rs[t2];
          const c64 x00 = Ac[p1 * DLD + p2], x01 = Ac[p1 * DLD + q2];
          const c64 x10 = Ac[q1 * DLD + p2], x11 = Ac[q1 * DLD + q2];
          // rows with rotation 1
          const c64 y00 = csub(cscale(c1, x00), cmul(s1, x10)), y01 = csub(cscale(c1, x01), cmul(s1, x11));
          const c64 y10 = cadd(cmul(ccj(s1), x00), cscale(c1, x10)), y11 = cadd(cmul(ccj(s1), x01), cscale(c1, x11));
          // columns with rotation 2
          An[p1 * DLD + p2] = csub(cscale(c2, y00), cmul(ccj(s2), y01));
          An[p1 * DLD + q2] = cadd(cmul(s2, y00), cscale(c2, y01));
          An[q1 * DLD + p2] = csub(cscale(c2, y10), cmul(ccj(s2), y11));
          An[q1 * DLD + q2] = cadd(cmul(s2, y10), cscale(c2, y11));
        } else {
          const int e = idx - nblk;
          const int j = e / npair, t = e - (e / npair) * npair;
          const int p = rp[t], q = rq[t];
          const double c = rc[t];
          const c64 sv = rs[t];
          const c64 vp = V[j * DLD + p], vq = V[j * DLD + q];
          V[j * DLD + p] = csub(cscale(c, vp), cmul(ccj(sv), vq));
          V[j * DLD + q] = cadd(cmul(sv, vp), cscale(c, vq));
        }
      }
      __syncthreads();
      c64* tmp = Ac;
      Ac = An;
      An = tmp;
    }
    if (rot[sl] == 0 || big[sl] == 0) break;
  }
  return Ac;
}

__device__ __forceinline__ int shp(const int* shape, int f, int ninst, int i) { return shape[f * ninst + i]; }

// ----------------------------------------------------------------------------------------
// kernels
// ----------------------------------------------------------------------------------------

constexpr int RBLK = 512;  // Rayleigh-Ritz block: one thread per Jacobi item up to 24 x 24

__global__ __launch_bounds__(RBLK) void k_dense_rr(DenseLaunch dl, LState* st, int* shape, const c64* gp,
                                                   double* Cx, double* lam, int diag_mode) {
  __shared__ c64 A[DMAX * DLD];
  __shared__ c64 A2[DMAX * DLD];
  __shared__ c64 V[DMAX * DLD];
  __shared__ int rank[DMAX];
  const int i = blockIdx.x, ninst = dl.ninst, pm = dl.pm;
  const int nact = shp(shape, F_NACT, ninst, i);
  if (nact == 0) return;
  const int np = shp(shape, F_NP, ninst, i);
  const int py = diag_mode ? nact : 2 * nact + np;
  const long esz = long(pm) * pm;
  for (int e = threadIdx.x; e < py * py; e += blockDim.x) {
    const int a = e / py, b = e - (e / py) * py;
    double sx = 0, sy = 0;
    const c64* p = gp + long(i) * dl.nchunk * esz + a * pm + b;
    for (int c = 0; c < dl.nchunk; ++c) {
      sx += p[c * esz].x;
      sy += p[c * esz].y;
    }
    A[a * DLD + b] = c64{sx, sy};
  }
  __syncthreads();
  // Hermitian part.  The Gram holds the 16 x 16 tiles on and above the tile diagonal:
  // entries in diagonal tiles are averaged with their mirror, the others are mirrored.
  for (int e = threadIdx.x; e < py * py; e += blockDim.x) {
    const int a = e / py, b = e - (e / py) * py;
    if (a < b) {
      const c64 u = A[a * DLD + b];
      c64 h = u;
      if ((a >> 4) == (b >> 4)) {
        const c64 v = A[b * DLD + a];
        h = c64{0.5 * (u.x + v.x), 0.5 * (u.y - v.y)};
      }
      A[a * DLD + b] = h;
      A[b * DLD + a] = ccj(h);
    } else if (a == b) {
      A[a * DLD + a].y = 0.0;
    }
  }
  __syncthreads();
  double* Ci = Cx + 2 * long(i) * esz;
  double* li = lam + long(i) * pm;
  if (diag_mode) {
    for (int e = threadIdx.x; e < py * nact; e += blockDim.x) {
      const int a = e / nact, q = e - (e / nact) * nact;
      Ci[2 * (a * pm + q)] = (a == q) ? 1.0 : 0.0;
      Ci[2 * (a * pm + q) + 1] = 0.0;
    }
    for (int q = threadIdx.x; q < nact; q += blockDim.x) li[q] = A[q * DLD + q].x;
    return;
  }
  const c64* Ad = d_jacobi(A, A2, V, py, dl.jac_stop);
  for (int j = threadIdx.x; j < py; j += blockDim.x) {
    const double wj = Ad[j * DLD + j].x;
    int r = 0;
    for (int l = 0; l < py; ++l) {
      const double wl = Ad[l * DLD + l].x;
      if (wl < wj || (wl == wj && l < j)) ++r;
    }
    rank[j] = r;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < py * py; e += blockDim.x) {
    const int a = e / py, j = e - (e / py) * py;
    const int q = rank[j];
    if (q < nact) {
      Ci[2 * (a * pm + q)] = V[a * DLD + j].x;
      Ci[2 * (a * pm + q) + 1] = V[a * DLD + j].y;
    }
  }
  for (int j = threadIdx.x; j < py; j += blockDim.x)
    if (rank[j] < nact) li[rank[j]] = Ad[j * DLD + j].x;
}

__global__ __launch_bounds__(DBLK) void k_dense_lock(DenseLaunch dl, LState* st, int* shape, const double* npart,
                                                     const double* Cx, double* Cp, double* rhist) {
  __shared__ c64 cX[DMAX * SLD];
  __shared__ c64 cP[DMAX * SLD];
  __shared__ c64 B[16 * SLD];
  __shared__ c64 O[16 * SLD];
  __shared__ c64 W[16 * SLD];
  __shared__ c64 Wr[DMAX * SLD];
  __shared__ double nrm[16];
  __shared__ int s_newly, s_finished;
  const int i = blockIdx.x, ninst = dl.ninst, pm = dl.pm;
  LState& s = st[i];
  const int nact = shp(shape, F_NACT, ninst, i);
  if (nact == 0) return;
  // one read of the state (its fields are dependent loads otherwise, each after the
  // stores that might alias it); thread 0 stores the fields it changes
  const LState s0 = s;
  // residual norms from the per-block partials: lane l sums blocks l, l + 64, ... of every
  // column (independent loads), then one wavefront reduction per column (fixed order)
  {
    double part[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) part[q] = 0.0;
    for (int b = threadIdx.x; b < dl.nblk; b += DBLK) {
      // every column's load issued before the sums (columns past the row's qmax re-read its
      // last entry and are not summed)
      const double* np = npart + (long(i) * dl.nblk + b) * dl.qmax;
      double v[16];
#pragma unroll
      for (int q = 0; q < 16; ++q) v[q] = np[q < dl.qmax ? q : dl.qmax - 1];
#pragma unroll
      for (int q = 0; q < 16; ++q)
        if (q < nact) part[q] += v[q];
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      double v = part[q];
#pragma unroll
      for (int off = DBLK / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, DBLK);
      if (threadIdx.x == 0 && q < nact) nrm[q] = sqrt(v);
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int M = s0.M;
    const int prev = s0.prev_nlocked;
    if (s0.niter <= dl.maxiter)
      for (int q = 0; q < nact; ++q) rhist[(long(i) * (dl.maxiter + 1) + s0.niter) * 16 + prev + q] = nrm[q];
    int nl = prev;
    if (s0.niter >= 1) {
      for (int q = 0; q < nact; ++q) {
        if (nrm[q] < s0.tol)
          ++nl;
        else
          break;
      }
    }
    const int conv = nl >= s0.ncheck ? 1 : 0;
    const int finished = conv || s0.niter >= dl.maxiter;
    s.nlocked = nl;
    s.newly_locked = nl - prev;
    s.converged = conv;
    if (finished && !conv) s.status = OMX_ERR_NOT_CONVERGED;
    s.finished = finished;
    shape[F_NLOCK * ninst + i] = nl;
    shape[F_NEWLY * ninst + i] = nl - prev;
    shape[F_NNEW * ninst + i] = finished ? 0 : M - nl;
    shape[F_NCOPY * ninst + i] = finished ? nact : nl - prev;
    shape[F_NPN * ninst + i] = (finished || s0.niter == 0) ? 0 : M - nl;
    s_newly = nl - prev;
    s_finished = finished;
  }
  __syncthreads();
  const int newly = s_newly;
  if (s_finished || s0.niter == 0) return;
  // cP = (cX - e)[:, xn], xn = newly .. nact-1 (dftk.jl:426-438)
  const int np = s0.np;
  const int py = 2 * nact + np;
  const int nn = nact - newly;
  const long esz = long(pm) * pm;
  const double* Ci = Cx + 2 * long(i) * esz;
  for (int e = threadIdx.x; e < py * nact; e += blockDim.x) {
    const int a = e / nact, q = e - (e / nact) * nact;
    cX[a * SLD + q] = c64{Ci[2 * (a * pm + q)], Ci[2 * (a * pm + q) + 1]};
  }
  __syncthreads();
  for (int e = threadIdx.x; e < py * nn; e += blockDim.x) {
    const int a = e / nn, jp = e - (e / nn) * nn;
    const int col = newly + jp;
    c64 v = cX[a * SLD + col];
    if (col < nn && a == newly + col) v.x -= 1.0;
    cP[a * SLD + jp] = v;
  }
  __syncthreads();
  const int rc = d_ortho_against(cP, py, nn, cX, nact, s0.eps, s0.ortho_tol, B, O, W, Wr,
                                 uint64_t(i) * 1000003ull + uint64_t(s0.niter));
  if (rc != 0 && threadIdx.x == 0) s.status = rc;
  double* Pi = Cp + 2 * long(i) * esz;
  for (int e = threadIdx.x; e < py * nn; e += blockDim.x) {
    const int a = e / nn, q = e - (e / nn) * nn;
    Pi[2 * (a * pm + q)] = cP[a * SLD + q].x;
    Pi[2 * (a * pm + q) + 1] = cP[a * SLD + q].y;
  }
}

__global__ __launch_bounds__(DBLK) void k_dense_chol(DenseLaunch dl, LState* st, const int* shape, int cnt_field,
                                                     const c64* gp, double* C) {
  __shared__ c64 O[16 * SLD];
  __shared__ c64 O0[16 * SLD];
  __shared__ c64 W[16 * SLD];
  const int i = blockIdx.x, ninst = dl.ninst, pm = dl.pm;
  const int n = shp(shape, cnt_field, ninst, i);
  if (n == 0) return;
  const long esz = long(pm) * pm;
  for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
    const int a = e / n, b = e - (e / n) * n;
    double sx = 0, sy = 0;
    const c64* p = gp + long(i) * dl.nchunk * esz + a * pm + b;
    for (int c = 0; c < dl.nchunk; ++c) {
      sx += p[c * esz].x;
      sy += p[c * esz].y;
    }
    O0[a * SLD + b] = c64{sx, sy};
  }
  __syncthreads();
  for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
    const int a = e / n, b = e - (e / n) * n;
    const c64 u = O0[a * SLD + b], v = O0[b * SLD + a];
    O[a * SLD + b] = c64{0.5 * (u.x + v.x), 0.5 * (u.y - v.y)};
  }
  __syncthreads();
  double tr = 0;
  for (int a = 0; a < n; ++a) tr += O[a * SLD + a].x;
  for (int e = threadIdx.x; e < n * n; e += blockDim.x) O0[(e / n) * SLD + e % n] = O[(e / n) * SLD + e % n];
  __syncthreads();
  bool ok = d_chol(O, SLD, n);
  double alpha = 100.0;
  int nbad = 0;
  while (!ok && nbad <= 10) {
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
      const int a = e / n, b = e - (e / n) * n;
      O[a * SLD + b] = O0[a * SLD + b];
      if (a == b) O[a * SLD + b].x += alpha * st[i].eps * tr;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < n; e += blockDim.x) O0[e * SLD + e].x += alpha * st[i].eps * tr;
    __syncthreads();
    alpha *= 10;
    ok = d_chol(O, SLD, n);
    ++nbad;
  }
  double* Ci = C + 2 * long(i) * esz;
  if (!ok) {
    if (threadIdx.x == 0) st[i].status = OMX_ERR_CHOLESKY;
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
      const int a = e / n, b = e - (e / n) * n;
      Ci[2 * (a * pm + b)] = a == b ? 1.0 : 0.0;
      Ci[2 * (a * pm + b) + 1] = 0.0;
    }
    return;
  }
  d_triinv(O, SLD, W, SLD, n);
  for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
    const int a = e / n, b = e - (e / n) * n;
    Ci[2 * (a * pm + b)] = W[a * SLD + b].x;
    Ci[2 * (a * pm + b) + 1] = W[a * SLD + b].y;
  }
}

// One pass of R <- (R - Z B) U^{-1}: from the Gram of [R Z]^H R (R = F_NNEW columns, Z =
// the M X columns plus F_NPN P columns) form B = Z^H R, S = R^H R - B^H B (the Gram of the
// projected R for orthonormal Z), the upper Cholesky S = U^H U with DFTK's shifting on
// failure (dftk.jl:163-192), and the coefficients C = [U^{-1}; -B U^{-1}] of the update
// R <- [R Z] C.
__global__ __launch_bounds__(DBLK) void k_dense_orth(DenseLaunch dl, LState* st, const int* shape, const c64* gp,
                                                     int pmg, double* C) {
  __shared__ c64 G[DMAX * SLD];
  __shared__ c64 S[16 * SLD];
  __shared__ c64 S0[16 * SLD];
  __shared__ c64 W[16 * SLD];
  const int i = blockIdx.x, ninst = dl.ninst, pm = dl.pm;
  const int nr = shp(shape, F_NNEW, ninst, i);
  if (nr == 0) return;
  const int nz = shp(shape, F_M, ninst, i) + shp(shape, F_NPN, ninst, i);
  const int pa = nz + nr;
  const long esz = long(pmg) * pmg;
  for (int e = threadIdx.x; e < pa * nr; e += blockDim.x) {
    const int a = e / nr, b = e - (e / nr) * nr;
    double sx = 0, sy = 0;
    const c64* p = gp + long(i) * dl.nchunk * esz + a * pmg + b;
    for (int c = 0; c < dl.nchunk; ++c) {
      sx += p[c * esz].x;
      sy += p[c * esz].y;
    }
    G[a * SLD + b] = c64{sx, sy};
  }
  __syncthreads();
  // G rows: 0 .. nr-1 = R^H R, nr .. nr+nz-1 = Z^H R = B
  for (int e = threadIdx.x; e < nr * nr; e += blockDim.x) {
    const int a = e / nr, b = e - (e / nr) * nr;
    c64 v = G[a * SLD + b];
    for (int l = 0; l < nz; ++l) v = csub(v, cmul(ccj(G[(nr + l) * SLD + a]), G[(nr + l) * SLD + b]));
    S0[a * SLD + b] = v;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < nr * nr; e += blockDim.x) {
    const int a = e / nr, b = e - (e / nr) * nr;
    const c64 u = S0[a * SLD + b], v = S0[b * SLD + a];
    S[a * SLD + b] = c64{0.5 * (u.x + v.x), 0.5 * (u.y - v.y)};
  }
  __syncthreads();
  double tr = 0;
  for (int a = 0; a < nr; ++a) tr += S[a * SLD + a].x;
  for (int e = threadIdx.x; e < nr * nr; e += blockDim.x) S0[(e / nr) * SLD + e % nr] = S[(e / nr) * SLD + e % nr];
  __syncthreads();
  const double eps = st[i].eps;
  bool ok = d_chol(S, SLD, nr);
  double alpha = 100.0;
  int nbad = 0;
  while (!ok && nbad <= 10) {
    for (int e = threadIdx.x; e < nr * nr; e += blockDim.x) {
      const int a = e / nr, b = e - (e / nr) * nr;
      S[a * SLD + b] = S0[a * SLD + b];
      if (a == b) S[a * SLD + b].x += alpha * eps * tr;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < nr; e += blockDim.x) S0[e * SLD + e].x += alpha * eps * tr;
    __syncthreads();
    alpha *= 10;
    ok = d_chol(S, SLD, nr);
    ++nbad;
  }
  double* Ci = C + 2 * long(i) * pm * pm;
  if (!ok) {
    if (threadIdx.x == 0) st[i].status = OMX_ERR_CHOLESKY;
    for (int e = threadIdx.x; e < pa * nr; e += blockDim.x) {
      const int a = e / nr, b = e - (e / nr) * nr;
      Ci[2 * (a * pm + b)] = (a == b) ? 1.0 : 0.0;
      Ci[2 * (a * pm + b) + 1] = 0.0;
    }
    return;
  }
  d_triinv(S, SLD, W, SLD, nr);
  for (int e = threadIdx.x; e < pa * nr; e += blockDim.x) {
    const int a = e / nr, q = e - (e / nr) * nr;
    c64 v;
    if (a < nr) {
      v = W[a * SLD + q];
    } else {
      v = c64{0, 0};
      for (int l = 0; l < nr; ++l) v = csub(v, cmul(G[a * SLD + l], W[l * SLD + q]));
    }
    Ci[2 * (a * pm + q)] = v.x;
    Ci[2 * (a * pm + q) + 1] = v.y;
  }
}

__global__ void k_advance(DenseLaunch dl, LState* st, int* shape) {
  const int ninst = dl.ninst;
  for (int i = threadIdx.x; i < ninst; i += blockDim.x) {
    LState& s = st[i];
    if (shape[F_NACT * ninst + i] == 0) continue;
    const int nnew = shape[F_NNEW * ninst + i];
    const int niter_before = s.niter;
    if (nnew == 0) {
      // finished this iteration (both X buffers already hold the final block): idle from
      // now on, so that later iterations of the batch leave this instance untouched
      shape[F_NACT * ninst + i] = 0;
      shape[F_NP * ninst + i] = 0;
      shape[F_NCOPY * ninst + i] = 0;
      continue;
    }
    s.prev_nlocked = s.nlocked;
    s.nact = nnew;
    s.np = niter_before >= 1 ? nnew : 0;
    s.niter = niter_before + 1;
    shape[F_XOFF * ninst + i] = s.nlocked;
    shape[F_NACT * ninst + i] = nnew;
    shape[F_ROFF * ninst + i] = shape[F_NEWLY * ninst + i];
    shape[F_NP * ninst + i] = s.np;
  }
}

__global__ void k_lobpcg_init(DenseLaunch dl, LState* st, int* shape, const int* active, const int* band) {
  const int ninst = dl.ninst;
  for (int i = threadIdx.x; i < ninst; i += blockDim.x) {
    LState& s = st[i];
    const int on = active[i];
    s.active = on;
    s.niter = 0;
    s.nlocked = s.prev_nlocked = 0;
    s.newly_locked = 0;
    s.nact = s.M;
    s.np = 0;
    s.converged = 0;
    s.finished = on ? 0 : 1;
    s.status = 0;
    shape[F_XOFF * ninst + i] = 0;
    shape[F_NACT * ninst + i] = on ? s.M : 0;
    shape[F_ROFF * ninst + i] = 0;
    shape[F_NP * ninst + i] = 0;
    shape[F_NLOCK * ninst + i] = 0;
    shape[F_NNEW * ninst + i] = on ? s.M : 0;
    shape[F_NEWLY * ninst + i] = 0;
    shape[F_NCOPY * ninst + i] = 0;
    shape[F_NPN * ninst + i] = 0;
    shape[F_M * ninst + i] = on ? s.M : 0;
    shape[F_ONE * ninst + i] = on ? 1 : 0;
    shape[F_EIG * ninst + i] = band[i];
  }
}

__global__ void k_newton(int ninst, int mper, NewtonState* ns, LState* st, const c64* rq, int rq_nchunk, int rq_pm,
                         const double* hp, int hnblk, double* kinst) {
  for (int i = threadIdx.x; i < ninst; i += blockDim.x) {
    NewtonState& n = ns[i];
    if (n.done || !st[i].active) continue;
    double lam = 0;
    const long esz = long(rq_pm) * rq_pm;
    for (int c = 0; c < rq_nchunk; ++c) lam += rq[(long(i) * rq_nchunk + c) * esz].x;
    double h = 0;
    const long v = long(i) * mper + n.band;
    for (int b = 0; b < hnblk; ++b) h += hp[v * hnblk + b];
    const double f = lam - n.omega2;
    const double slope = 2.0 * h;
    n.lam = lam;
    n.f = f;
    n.slope = slope;
    n.nsolves += 1;
    if (fabs(f) <= n.k_tol) {
      n.done = 1;
    } else {
      const double knew = n.k - f / slope;
      if (!isfinite(knew)) {
        n.done = 2;
      } else if (fabs(knew - n.k) <= 4.0 * 2.220446049250313e-16 * fmax(fabs(n.k), 1.0)) {
        n.done = 1;
      } else if (n.nsolves >= n.max_solves) {
        n.done = 2;
      } else {
        n.k = knew;
      }
    }
    kinst[i] = n.k;
  }
}

}  // namespace

void launch_dense_rr(const DenseLaunch& dl, LState* st, int* shape, const c64* gram_partial, double* Cx,
                     double* lam, int diag_mode, hipStream_t stream) {
  hipLaunchKernelGGL(k_dense_rr, dim3(dl.ninst), dim3(RBLK), 0, stream, dl, st, shape, gram_partial, Cx, lam,
                     diag_mode);
  OMX_HIP_CHECK(hipGetLastError());
}

void launch_dense_lock(const DenseLaunch& dl, LState* st, int* shape, const double* norm_partial, const double* Cx,
                       double* Cp, double* resid_hist, hipStream_t stream) {
  hipLaunchKernelGGL(k_dense_lock, dim3(dl.ninst), dim3(DBLK), 0, stream, dl, st, shape, norm_partial, Cx, Cp,
                     resid_hist);
  OMX_HIP_CHECK(hipGetLastError());
}

void launch_dense_chol(const DenseLaunch& dl, LState* st, const int* shape, int cnt_field, const c64* gram_partial,
                       double* C, hipStream_t stream) {
  hipLaunchKernelGGL(k_dense_chol, dim3(dl.ninst), dim3(DBLK), 0, stream, dl, st, shape, cnt_field, gram_partial,
                     C);
  OMX_HIP_CHECK(hipGetLastError());
}

void launch_dense_orth(const DenseLaunch& dl, LState* st, const int* shape, const c64* gram_partial, int pmg,
                       double* C, hipStream_t stream) {
  hipLaunchKernelGGL(k_dense_orth, dim3(dl.ninst), dim3(DBLK), 0, stream, dl, st, shape, gram_partial, pmg, C);
  OMX_HIP_CHECK(hipGetLastError());
}

void launch_advance(const DenseLaunch& dl, LState* st, int* shape, hipStream_t stream) {
  hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, stream, dl, st, shape);
  OMX_HIP_CHECK(hipGetLastError());
}

void launch_lobpcg_init(const DenseLaunch& dl, LState* st, int* shape, const int* active, const int* band,
                        hipStream_t stream) {
  hipLaunchKernelGGL(k_lobpcg_init, dim3(1), dim3(64), 0, stream, dl, st, shape, active, band);
  OMX_HIP_CHECK(hipGetLastError());
}

void launch_newton(int ninst, int mper, NewtonState* ns, LState* st, const c64* rq_partial, int rq_nchunk, int rq_pm,
                   const double* hmkh_partial, int hmkh_nblk, double* kinst, hipStream_t stream) {
  hipLaunchKernelGGL(k_newton, dim3(1), dim3(64), 0, stream, ninst, mper, ns, st, rq_partial, rq_nchunk, rq_pm,
                     hmkh_partial, hmkh_nblk, kinst);
  OMX_HIP_CHECK(hipGetLastError());
}

}  // namespace omx
