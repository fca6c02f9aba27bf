// nspt_oracle.cpp -- CPU restatement of the PRlgt2020 NSPT Langevin step.
//
// TEST INFRASTRUCTURE ONLY.  This file is the oracle that the HIP product path
// is checked against.  Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may load it; the product library never links
// or calls it.
//
// It restates, in plain scalar C++, the arithmetic of the reference:
//   * series algebra of ptSU3 / ptSU2          QCDpt.h:92-114, 176-220; QCDpt.cc:37-209, 318-537
//   * 3x3 / 2x2 complex products, dag, Tr      MyMath.h:791-849, 1001, 1064-1080
//   * periodic lattice indexing + neighbours   lattice.h:37, 139-233, 1194-1222
//   * twisted boundary conditions (as the      nspt.cc:97-256, lattice.h:156-1053,
//     conjugation W U W^dag of links read      twistmatrices.h
//     across a twisted boundary)
//   * staple                                   QCDenvNODEpt_test.h:554-557
//   * gauge_wilson (Jacobi and on-the-fly)     nspt.cc:258-600
//   * zero_modes_subtraction                   nspt.cc:605-654
//   * stochastic_gauge_fixing (in place)       nspt.cc:660-862
//   * FAstochastic_gauge_fixing                nspt.cc:866-948 (FFTW plan_many semantics of
//                                              Quenched.cc:827-838 restated as a direct DFT)
//   * plaquette_measure (OpenMP normalisation) nspt.cc:1316-1406
//   * ComputeLoops (1x2, 2x2, 2x4, 4x4 loops)  nspt.cc:1156-1312
//   * knorm (__K_MOM_ANALYSIS__)               nspt.cc:919-944, Quenched.cc:792-815
//   * noise: SU3rand / SU2rand                 MyMath.cc:256-368
//   * RAND55 + Box-Muller                      MyRand.h:81-108, MyRand.cc:73-130
//   * ranlxd (portable path)                   ranlxd/ranlxd.cc:332-601
//   * RK2 (absent from the reference): second-order Runge-Kutta of Batrouni,
//     Katz, Kronfeld, Lepage, Svetitsky, Wilson, PRD 32 (1985) 2736, with the
//     same noise in predictor and corrector; the (1 + C_A eps_raw/6) factor
//     becomes a two-order shift in the beta^-1/2 expansion.
//   * Philox4x32-10 device-noise restatement (Salmon et al., SC'11).
//
// Pinning of each part (tests/test_oracle_*.py): golden vectors written by
// oracle/_ref (the reference's own algebra + RNG sources compiled here), the
// ranlxd known-answer test of ranlxd/test/test.cc:47-102, numpy's FFT for the
// DFT, and closed forms / invariants (log(exp X) = X, unitarity, gauge
// covariance, cold-start first step, brute-force plaquette); the twisted
// lattice against the periodic one on the unwrapped covering lattice.
//
// Series layout: (1 + nc*nc*N) complex, index 0 = flag (order-0 scalar),
// 1 + (k-1)*nc*nc + r*nc + c = element (r,c) of order k.  This is the memory
// layout of ptSU3 {Cplx flag; SU3 ptU[allocORD];} with allocORD = N.
// Field layout: link (site, mu) at series index 4*site + mu (QCDenvNODEpt_test.h:521).

#include <chrono>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

// Scalar type of the series arithmetic: double, or (built with
// -DORC_EXTENDED=1) the x87 80-bit long double, whose 64-bit significand
// gives an accuracy reference for FP64 steps at high orders.
#if ORC_EXTENDED
using real = long double;
#else
using real = double;
#endif
constexpr real kPi = 3.14159265358979323846264338327950288L;

struct cx {
  real re, im;
};
inline cx mk(real r, real i) { return cx{r, i}; }
inline cx operator+(cx a, cx b) { return cx{a.re + b.re, a.im + b.im}; }
inline cx operator-(cx a, cx b) { return cx{a.re - b.re, a.im - b.im}; }
inline cx operator*(cx a, cx b) { return cx{a.re * b.re - a.im * b.im, a.im * b.re + a.re * b.im}; }
inline cx operator*(real x, cx b) { return cx{x * b.re, x * b.im}; }
inline cx& operator+=(cx& a, cx b) { a.re += b.re; a.im += b.im; return a; }
inline cx& operator-=(cx& a, cx b) { a.re -= b.re; a.im -= b.im; return a; }
inline cx conj(cx a) { return cx{a.re, -a.im}; }
inline bool is_zero(cx a) { return a.re == 0.0 && a.im == 0.0; }

// ---------------------------------------------------------------- matrices
// Row-major nc x nc complex matrices (MyMath.h:1697-1706).

// C = A * B; each entry summed over m in ascending order (MyMath.h:791-849).
inline void mat_mul(int nc, const cx* A, const cx* B, cx* C) {
  for (int r = 0; r < nc; r++)
    for (int c = 0; c < nc; c++) {
      real re = 0, im = 0;
      for (int m = 0; m < nc; m++) {
        const cx a = A[r * nc + m], b = B[m * nc + c];
        re += a.re * b.re - a.im * b.im;
        im += a.re * b.im + a.im * b.re;
      }
      C[r * nc + c] = cx{re, im};
    }
}

// C += A * B
inline void mat_mac(int nc, const cx* A, const cx* B, cx* C) {
  cx t[9];
  mat_mul(nc, A, B, t);
  for (int e = 0; e < nc * nc; e++) C[e] += t[e];
}

inline void mat_dag(int nc, const cx* A, cx* B) {
  for (int r = 0; r < nc; r++)
    for (int c = 0; c < nc; c++) B[c * nc + r] = conj(A[r * nc + c]);
}

inline cx mat_tr(int nc, const cx* A) {
  cx t = A[0];
  for (int r = 1; r < nc; r++) t = t + A[r * nc + r];
  return t;
}

// ---------------------------------------------------------------- series
struct Ser {
  int nc, N;
  std::vector<cx> v;  // 1 + nc*nc*N
  Ser() : nc(0), N(0) {}
  Ser(int nc_, int N_, cx flag = cx{0, 0}) : nc(nc_), N(N_), v(1 + nc_ * nc_ * N_, cx{0, 0}) { v[0] = flag; }
  cx& flag() { return v[0]; }
  const cx& flag() const { return v[0]; }
  cx* ord(int k) { return &v[1 + (k - 1) * nc * nc]; }              // k = 1..N
  const cx* ord(int k) const { return &v[1 + (k - 1) * nc * nc]; }
  int ne() const { return nc * nc; }
};

// Truncated Cauchy product L*R (QCDpt.h:92-114): orders i+j <= N from the
// matrix parts, then the scalar-flag terms R.flag*L_k + L.flag*R_k.
Ser ser_mul(const Ser& L, const Ser& R) {
  const int nc = L.nc, N = L.N;
  Ser C(nc, N);
  for (int i = 1; i <= N - 1; i++)         // order of R
    for (int j = 1; j <= N - i; j++)       // order of L
      mat_mac(nc, L.ord(j), R.ord(i), C.ord(i + j));
  if (is_zero(L.flag()) && is_zero(R.flag())) {
    C.flag() = cx{0, 0};
    return C;
  }
  C.flag() = L.flag() * R.flag();
  for (int k = 1; k <= N; k++) {
    cx* c = C.ord(k);
    const cx* l = L.ord(k);
    const cx* r = R.ord(k);
    for (int e = 0; e < C.ne(); e++) c[e] += R.flag() * l[e] + L.flag() * r[e];
  }
  return C;
}

Ser ser_add(const Ser& A, const Ser& B) {
  Ser C = A;
  for (size_t i = 0; i < C.v.size(); i++) C.v[i] += B.v[i];
  return C;
}
Ser ser_sub(const Ser& A, const Ser& B) {
  Ser C = A;
  for (size_t i = 0; i < C.v.size(); i++) C.v[i] -= B.v[i];
  return C;
}
void ser_scale(Ser& A, real x) {
  for (auto& z : A.v) z = x * z;
}
void ser_div(Ser& A, real x) {
  for (auto& z : A.v) z = cx{z.re / x, z.im / x};
}

Ser ser_dag(const Ser& A) {
  Ser B(A.nc, A.N);
  for (int k = 1; k <= A.N; k++) mat_dag(A.nc, A.ord(k), B.ord(k));
  B.flag() = conj(A.flag());
  return B;
}

// reH (QCDpt.cc:197-209, SU(2): 510-521): per order (M - M^dag)/2 minus Tr/nc.
void ser_reH(Ser& A) {
  const int nc = A.nc;
  const real inv = (nc == 3) ? real(1) / 3 : real(1) / 2;
  for (int k = 1; k <= A.N; k++) {
    cx* m = A.ord(k);
    cx d[9];
    mat_dag(nc, m, d);
    for (int e = 0; e < nc * nc; e++) m[e] = 0.5 * (m[e] - d[e]);
    const cx tr = inv * mat_tr(nc, m);
    for (int r = 0; r < nc; r++) m[r * nc + r] -= tr;
  }
  A.flag() = cx{0, 0};
}

// Trless (QCDpt.cc:124-136)
void ser_trless(Ser& A) {
  const int nc = A.nc;
  const real inv = (nc == 3) ? real(1) / 3 : real(1) / 2;
  for (int k = 1; k <= A.N; k++) {
    cx* m = A.ord(k);
    const cx z = inv * mat_tr(nc, m);
    for (int r = 0; r < nc; r++) m[r * nc + r] -= z;
  }
  A.flag() = cx{0, 0};
}

// Tr (QCDpt.cc:115-122): tt[0] = nc*flag, tt[k] = Tr of order k.
void ser_tr(const Ser& A, cx* tt) {
  for (int k = 1; k <= A.N; k++) tt[k] = mat_tr(A.nc, A.ord(k));
  tt[0] = (real)A.nc * A.flag();
}

// Powers step shared by exp and log: Bnew_{iA+iB} += B_iB * A_iA for iB >= i-1
// (QCDpt.h:181-187 and 204-211).
void power_step(const Ser& B, const Ser& A, int i, Ser& Bnew) {
  const int N = A.N;
  for (auto& z : Bnew.v) z = cx{0, 0};
  for (int iA = 1; iA <= N - 1; iA++)
    for (int iB = i - 1; iB <= N - iA; iB++) mat_mac(A.nc, B.ord(iB), A.ord(iA), Bnew.ord(iA + iB));
}

// exp (QCDpt.h:200-220): sum_k A^k/k! on the matrix part; result flag = 1.
Ser ser_exp(const Ser& A) {
  Ser B = A, Bnew(A.nc, A.N), res = A;
  for (int i = 2; i <= A.N; i++) {
    power_step(B, A, i, Bnew);
    const real den = real(1) / i;
    B = Bnew;
    ser_scale(B, den);
    for (size_t q = 0; q < res.v.size(); q++) res.v[q] += B.v[q];
  }
  res.flag() = cx{1, 0};
  return res;
}

// log (QCDpt.h:176-198): sum (-1)^{k+1} X^k/k with X = matrix part of U
// (flag assumed 1); result flag = 0.
Ser ser_log(const Ser& U) {
  Ser B = U, Bnew(U.nc, U.N), res = U;
  B.flag() = cx{0, 0};
  real segno = -1;
  for (int i = 2; i <= U.N; i++) {
    power_step(B, U, i, Bnew);
    for (int q = 1; q < (int)B.v.size(); q++) B.v[q] = Bnew.v[q];
    const real aux = segno / i;
    for (size_t q = 0; q < res.v.size(); q++) res.v[q] += aux * B.v[q];
    segno = -segno;
  }
  res.flag() = cx{0, 0};
  return res;
}

// ---------------------------------------------------------------- twist
// Twisted boundary conditions (TWISTED_BC, choices.h:43-55; nspt.cc:97-256;
// lattice.h:156-1053): the field obeys U(x + L_d e_d) = W_d U(x) W_d^dag for
// every twisted direction d, W_d the twist matrix of twistmatrices.h.  The
// reference stores twisted copies of the boundary links in halo sites
// (twist_boundary, nspt.cc:163-256: forward copies W A W^dag, backward copies
// W^dag A W); here a link reached across the boundary of d, n_d times
// (n_d = +-1, +-2), is conjugated by W_d^(n_d) when it is read.  Twist
// matrices of different directions commute up to a centre phase, so the
// order of the conjugations does not matter.
struct TwistState {
  int mask = 0;  // bit d: direction d twisted
  int nc = 0;
  cx om[4][9];
};
TwistState g_tw;

// Loops over links or sites run in parallel (OpenMP) on lattices large
// enough to pay for it; every sum is still formed in the serial order, so
// the results do not depend on it.
bool g_par = false;

// ---------------------------------------------------------------- lattice
struct Latt {
  int L[4];
  int V;
  std::vector<int> up, dn;  // [site*4 + mu]
  // Twisted lattices name the sites reached by neighbour steps with extended
  // indices e = site + V*code, code digit d (base 5) = n_d mod 5, n_d the
  // net number of crossings of the boundary of twisted direction d.
  int tw = 0;
  static constexpr int P5[4] = {1, 5, 25, 125};
  int fwd(int e, int d) const {
    if (!tw) return up[e * 4 + d];
    const int s = e % V;
    int code = e / V;
    int x[4];
    coords(s, x);
    if (((tw >> d) & 1) && x[d] == L[d] - 1) code = bump(code, d, 1);
    return up[s * 4 + d] + V * code;
  }
  int bwd(int e, int d) const {
    if (!tw) return dn[e * 4 + d];
    const int s = e % V;
    int code = e / V;
    int x[4];
    coords(s, x);
    if (((tw >> d) & 1) && x[d] == 0) code = bump(code, d, -1);
    return dn[s * 4 + d] + V * code;
  }
  static int bump(int code, int d, int by) {
    const int dig = (code / P5[d]) % 5;
    return code + (((dig + by + 5) % 5) - dig) * P5[d];
  }
  explicit Latt(const int* sz) {
    tw = g_tw.mask;
    for (int d = 0; d < 4; d++) L[d] = sz[d];
    V = L[0] * L[1] * L[2] * L[3];
    g_par = V >= 4096;
    up.resize(4 * V);
    dn.resize(4 * V);
    for (int s = 0; s < V; s++) {
      int x[4];
      coords(s, x);
      for (int d = 0; d < 4; d++) {
        int y[4] = {x[0], x[1], x[2], x[3]};
        y[d] = (x[d] + 1) % L[d];
        up[s * 4 + d] = index(y);
        y[d] = (x[d] - 1 + L[d]) % L[d];
        dn[s * 4 + d] = index(y);
      }
    }
  }
  // lattice.h:37 GETINDEX, x3 fastest
  int index(const int* x) const { return ((x[0] * L[1] + x[1]) * L[2] + x[2]) * L[3] + x[3]; }
  void coords(int s, int* x) const {
    x[3] = s % L[3];
    s /= L[3];
    x[2] = s % L[2];
    s /= L[2];
    x[1] = s % L[1];
    x[0] = s / L[1];
  }
};

// Conjugation of the twisted copy reached with crossing code `code`:
// M = prod_d W_d^(n_d) (W^-1 = W^dag).
void twist_matrix(int nc, int code, cx* M) {
  for (int q = 0; q < nc * nc; q++) M[q] = cx{(q / nc == q % nc) ? 1.0 : 0.0, 0.0};
  for (int d = 0; d < 4; d++) {
    int n = (code / Latt::P5[d]) % 5;
    if (n > 2) n -= 5;
    cx Wd[9], T[9];
    if (n < 0)
      mat_dag(nc, g_tw.om[d], Wd);
    else
      std::memcpy(Wd, g_tw.om[d], sizeof(cx) * nc * nc);
    for (int i = 0; i < (n < 0 ? -n : n); i++) {
      mat_mul(nc, M, Wd, T);
      std::memcpy(M, T, sizeof(cx) * nc * nc);
    }
  }
}

struct Field {
  int nc, N, S;
  const Latt* lat;
  real* data;  // interleaved complex, [(4*site+mu)*S + j]
  // e: a site or an extended index (twisted lattice): the twisted copy
  // M U M^dag of the stored link
  Ser get(int e, int mu) const {
    const int V = lat->V, site = e % V, code = e / V;
    Ser s(nc, N);
    std::memcpy(s.v.data(), data + 2 * (size_t)(4 * site + mu) * S, sizeof(cx) * S);
    if (code) {
      cx M[9], Md[9], T[9];
      twist_matrix(nc, code, M);
      mat_dag(nc, M, Md);
      for (int k = 1; k <= N; k++) {
        mat_mul(nc, M, s.ord(k), T);
        mat_mul(nc, T, Md, s.ord(k));
      }
    }
    return s;
  }
  // a twisted copy is written back as the stored link M^dag T M
  void set(int e, int mu, const Ser& s) {
    const int V = lat->V, site = e % V, code = e / V;
    real* dst = data + 2 * (size_t)(4 * site + mu) * S;
    std::memcpy(dst, s.v.data(), sizeof(cx) * S);
    if (code) {
      Ser u = s;
      cx M[9], Md[9], T[9];
      twist_matrix(nc, code, M);
      mat_dag(nc, M, Md);
      for (int k = 1; k <= N; k++) {
        mat_mul(nc, Md, s.ord(k), T);
        mat_mul(nc, T, M, u.ord(k));
      }
      std::memcpy(dst, u.v.data(), sizeof(cx) * S);
    }
  }
};

// staple(n, mu, nu) = U_nu(n+mu) [U_nu(n) U_mu(n+nu)]^dag
//                   + [U_mu(n-nu) U_nu(n-nu+mu)]^dag U_nu(n-nu)   (QCDenvNODEpt_test.h:554-557)
Ser staple(const Field& F, int n, int mu, int nu) {
  const Latt& Z = *F.lat;
  const int n_pmu = Z.fwd(n, mu), n_pnu = Z.fwd(n, nu), n_mnu = Z.bwd(n, nu);
  const int n_mnu_pmu = Z.fwd(n_mnu, mu);
  Ser a = ser_mul(F.get(n_pmu, nu), ser_dag(ser_mul(F.get(n, nu), F.get(n_pnu, mu))));
  Ser b = ser_mul(ser_dag(ser_mul(F.get(n_mnu, mu), F.get(n_mnu_pmu, nu))), F.get(n_mnu, nu));
  return ser_add(a, b);
}

// staple2x1(n, mu, nu) (QCDenvNODEpt_test.h:579-611): the six 5-link
// staples of the 2x1 and 1x2 rectangles through link (n, mu) in plane
// (mu, nu), with the reference's neighbour letters and product order.
Ser staple2x1(const Field& F, int n, int mu, int nu) {
  const Latt& Z = *F.lat;
  auto up = [&](int s, int d) { return Z.fwd(s, d); };
  auto dn = [&](int s, int d) { return Z.bwd(s, d); };
  auto U = [&](int s, int d) { return F.get(s, d); };
  int a = up(n, mu), b = up(a, nu), c = up(n, nu), d = up(c, nu), e = up(a, mu), f = dn(n, mu), g = up(f, nu);
  Ser t = ser_mul(ser_mul(U(a, nu), U(b, nu)), ser_dag(ser_mul(ser_mul(U(n, nu), U(c, nu)), U(d, mu))));
  t = ser_add(t, ser_mul(ser_mul(U(a, mu), U(e, nu)), ser_dag(ser_mul(ser_mul(U(n, nu), U(c, mu)), U(b, mu)))));
  t = ser_add(t, ser_mul(ser_mul(U(a, nu), ser_dag(ser_mul(ser_mul(U(f, nu), U(g, mu)), U(c, mu)))), U(f, mu)));
  b = dn(a, nu);
  c = dn(n, nu);
  d = dn(c, nu);
  const int e2 = dn(b, nu), f2 = up(b, mu), g2 = dn(c, mu), h = dn(n, mu);
  t = ser_add(t, ser_mul(ser_mul(ser_dag(ser_mul(ser_mul(U(d, mu), U(e2, nu)), U(b, nu))), U(d, nu)), U(c, nu)));
  t = ser_add(t, ser_mul(ser_mul(U(a, mu), ser_dag(ser_mul(ser_mul(U(c, mu), U(b, mu)), U(f2, nu)))), U(c, nu)));
  t = ser_add(t, ser_mul(ser_mul(ser_dag(ser_mul(ser_mul(U(g2, mu), U(c, mu)), U(b, nu))), U(g2, nu)), U(h, mu)));
  return t;
}

Ser staple_sum(const Field& F, int n, int mu) {
  Ser W(F.nc, F.N);  // zero(): flag 0
  for (int nu = 0; nu < 4; nu++)
    if (nu != mu) {
      Ser s = staple(F, n, mu, nu);
      for (size_t q = 0; q < W.v.size(); q++) W.v[q] += s.v[q];
    }
  return W;
}

// Observable accumulators of gauge_wilson (nspt.cc:392-402) and their
// normalisation (nspt.cc:570-597).
struct Obs {
  std::vector<cx> w1;             // N+1
  std::vector<real> norm;         // N
  std::vector<cx> trU;            // 4*N
  std::vector<cx> w2;             // N+1, improved actions only (nspt.cc:405-410)
  bool improved = false;
  void init(int N, bool impr = false) {
    w1.assign(N + 1, cx{0, 0});
    norm.assign(N, 0.0);
    trU.assign(4 * N, cx{0, 0});
    w2.assign(N + 1, cx{0, 0});
    improved = impr;
  }
  // One link's terms (w: Tr W per order; nrm, tr: Re Tr(U_k U_k^dag) and
  // Tr U_k; w2: Tr W2 of an improved action), recorded by a parallel loop
  // and added in link order afterwards, so that every sum runs in the
  // serial order.
  struct Rec {
    std::vector<cx> w, tr, w2;
    std::vector<real> nrm;
    int mu = 0;
  };
  static void record2(const Ser& W2, Rec& r) {
    r.w2.assign(W2.N + 1, cx{0, 0});
    ser_tr(W2, r.w2.data());
  }
  static void record(const Ser& U, const Ser& W, int mu, Rec& r) {
    const int N = U.N, nc = U.nc;
    r.mu = mu;
    r.w.assign(N + 1, cx{0, 0});
    ser_tr(W, r.w.data());
    r.nrm.assign(N, 0);
    r.tr.assign(N, cx{0, 0});
    for (int k = 1; k <= N; k++) {
      cx d[9] = {}, p[9] = {};
      mat_dag(nc, U.ord(k), d);
      mat_mul(nc, U.ord(k), d, p);
      r.nrm[k - 1] = mat_tr(nc, p).re;
      r.tr[k - 1] = mat_tr(nc, U.ord(k));
    }
  }
  void add(const Rec& r) {
    const int N = (int)norm.size();
    for (int k = 0; k <= N; k++) w1[k] += r.w[k];
    for (int k = 0; k < N; k++) {
      norm[k] += r.nrm[k];
      trU[r.mu * N + k] += r.tr[k];
    }
    if (!r.w2.empty())
      for (int k = 0; k <= N; k++) w2[k] += r.w2[k];
  }
  void accumulate2(const Ser& W2) {
    Rec r;
    record2(W2, r);
    for (int k = 0; k <= W2.N; k++) w2[k] += r.w2[k];
  }
  void accumulate(const Ser& U, const Ser& W, int mu) {
    Rec r;
    record(U, W, mu, r);
    add(r);
  }
  // out: plaq[2(N+1)], norm[N], trU[2*4N] normalised as written to file
  void write(int nc, int V, double* out) const {
    const int N = (int)norm.size();
    const real pn = (nc == 3) ? real(72) * V : real(48) * V;
    const real nn = (nc == 3) ? real(12) * V : real(8) * V;
    const real tn = (nc == 3) ? real(3) * V : real(2) * V;
    for (int k = 0; k <= N; k++) {
      out[2 * k] = w1[k].re / pn;
      out[2 * k + 1] = w1[k].im / pn;
    }
    double* o = out + 2 * (N + 1);
    for (int k = 0; k < N; k++) o[k] = norm[k] / nn;
    o += N;
    for (int q = 0; q < 4 * N; q++) {
      o[2 * q] = trU[q].re / tn;
      o[2 * q + 1] = trU[q].im / tn;
    }
    if (improved) {
      o += 8 * N;
      for (int k = 0; k <= N; k++) {
        o[2 * k] = w2[k].re / pn;
        o[2 * k + 1] = w2[k].im / pn;
      }
    }
  }
};

// Drift W = tau_g * reH(U * sum of staples), tau_g = -eps/(2 nc)
// (nspt.cc:383-417, Quenched.cc:584-589).  Optionally accumulates observables.
// c1 != 0: improved action (nspt.cc:385-412, Quenched.cc:313-322),
// W = tau_g reH(c0 U Sigma + c1 U Sigma_2x1) with c0 = 1 - 8 c1.
Ser drift(const Field& F, int n, int mu, real tau_g, Obs* obs, double c1 = 0.0, Obs::Rec* rec = nullptr) {
  Ser U = F.get(n, mu);
  Ser W = ser_mul(U, staple_sum(F, n, mu));
  if (obs) obs->accumulate(U, W, mu);
  if (rec) Obs::record(U, W, mu, *rec);
  if (c1 != 0.0) {
    Ser R(F.nc, F.N);
    for (int nu = 0; nu < 4; nu++)
      if (nu != mu) R = ser_add(R, staple2x1(F, n, mu, nu));
    Ser W2 = ser_mul(U, R);
    if (obs) obs->accumulate2(W2);
    if (rec) Obs::record2(W2, *rec);
    ser_scale(W, 1.0 - 8.0 * c1);
    ser_scale(W2, c1);
    W = ser_add(W, W2);
  }
  ser_reH(W);
  ser_scale(W, tau_g);
  return W;
}

void sub_noise(Ser& W, real stau, const double* eta) {
  cx* o1 = W.ord(1);
  for (int e = 0; e < W.ne(); e++) {
    o1[e].re -= stau * eta[2 * e];
    o1[e].im -= stau * eta[2 * e + 1];
  }
}

// ---------------------------------------------------------------- DFT
// Unnormalised multi-dimensional DFT with FFTW plan_many semantics:
// howmany components, element stride `stride`, distance 1 between
// components, sign -1 forward / +1 backward, row-major dims (x3 fastest).
void dft_many(const int* L, int howmany, int stride, cx* data, int sign) {
  const int V = L[0] * L[1] * L[2] * L[3];
  const int sstride[4] = {L[1] * L[2] * L[3], L[2] * L[3], L[3], 1};
  for (int d = 0; d < 4; d++) {
    const int n = L[d];
    std::vector<cx> tw(n);
    for (int j = 0; j < n; j++) {
      const real ang = 2 * kPi * j / n;
      tw[j] = cx{std::cos(ang), sign * std::sin(ang)};
    }
#pragma omp parallel for schedule(dynamic, 8) if (g_par)
    for (int s = 0; s < V; s++) {
      // visit each line once: sites with coordinate 0 along d
      if ((s / sstride[d]) % n != 0) continue;
      std::vector<cx> line(n), out(n);
      for (int comp = 0; comp < howmany; comp++) {
        for (int j = 0; j < n; j++) line[j] = data[(size_t)(s + j * sstride[d]) * stride + comp];
        for (int k = 0; k < n; k++) {
          cx acc{0, 0};
          for (int j = 0; j < n; j++) acc += line[j] * tw[(size_t)j * k % n];
          out[k] = acc;
        }
        for (int k = 0; k < n; k++) data[(size_t)(s + k * sstride[d]) * stride + comp] = out[k];
      }
    }
  }
}

#if !ORC_EXTENDED
// ---------------------------------------------------------------- RNGs
// ranlxd, portable path (ranlxd/ranlxd.cc:332-531): four interleaved 24-bit
// subtract-with-borrow lanes, 12 vectors of (low word, high word) x 4 lanes,
// lags (ir, jr = ir+7 mod 12), luxury p = 202 (level 1) or 397 (level 2).
struct Ranlxd {
  int num[96];      // vector v: low words at 8v+0..3, high words at 8v+4..7
  int carry[4];
  int pr, prm, ir, jr, is, is_old;
  int nxt[96];
  double one_bit;

  Ranlxd(int seed, int level) {
    one_bit = std::ldexp(1.0, -24);
    for (int k = 0; k < 96; k++) nxt[k] = (k % 4 == 3) ? (k + 5) % 96 : (k + 1) % 96;
    pr = (level == 2) ? 397 : 202;
    int xbit[31];
    int s = seed;
    for (int k = 0; k < 31; k++) {
      xbit[k] = s % 2;
      s /= 2;
    }
    int ibit = 0, jbit = 18;
    for (int lane = 0; lane < 4; lane++)
      for (int k = 0; k < 24; k++) {
        int ix = 0;
        for (int l = 0; l < 24; l++) {
          ix = 2 * ix + xbit[ibit];
          xbit[ibit] = (xbit[ibit] + xbit[jbit]) % 2;
          ibit = (ibit + 1) % 31;
          jbit = (jbit + 1) % 31;
        }
        if (k % 4 != lane) ix = 16777215 - ix;
        num[4 * k + lane] = ix;
      }
    for (int l = 0; l < 4; l++) carry[l] = 0;
    ir = 0;
    jr = 7;
    is = 91;
    is_old = 0;
    prm = pr % 12;
  }

  void swb(int vi, int vj) {
    const int base = 0x1000000, mask = 0xffffff;
    int* xi = &num[8 * vi];
    const int* xj = &num[8 * vj];
    for (int l = 0; l < 4; l++) {
      int d = xj[l] - xi[l] - carry[l];
      xi[4 + l] += (d < 0);
      xi[l] = (d + base) & mask;
    }
    for (int l = 0; l < 4; l++) {
      int d = xj[4 + l] - xi[4 + l];
      carry[l] = (d < 0);
      xi[4 + l] = (d + base) & mask;
    }
  }

  void update() {
    int vi = ir, vj = jr;
    for (int k = 0; k < pr; k++) {
      swb(vi, vj);
      vi = (vi + 1) % 12;
      vj = (vj + 1) % 12;
    }
    ir = (ir + prm) % 12;
    jr = (jr + prm) % 12;
    is = 8 * ir;
    is_old = is;
  }

  double next() {
    is = nxt[is];
    if (is == is_old) update();
    return one_bit * ((double)num[is + 4] + one_bit * (double)num[is]);
  }
};

// RAND55 additive lagged Fibonacci (MyRand.cc:73-110, MyRand.h:81-91) with
// the Box-Muller pairing of generate_gauss (MyRand.cc:115-130).  The signed
// wrap-around of the reference's seeding is reproduced with unsigned 64-bit
// arithmetic (same bits).
struct Rand55 {
  uint64_t ring[55];
  int nx[55];
  int n1, n2;
  bool gen;
  double r, t;
  explicit Rand55(long idum) {
    for (int i = 0; i < 54; i++) nx[i] = i + 1;
    nx[54] = 0;
    uint64_t tmp = (uint64_t)161803398L + (uint64_t)idum;
    ring[54] = tmp;
    uint64_t aux = 1;
    int j = 20;
    for (int i = 0; i < 54; i++, j += 21) {
      j %= 55;
      ring[j] = aux;
      aux += tmp;
      tmp = ring[j];
    }
    n1 = 0;
    n2 = 31;
    for (int q = 0; q < 55 * 4; q++) ranf();
    gen = false;
    r = t = 0.0;
  }
  double ranf() {
    n1 = nx[n1];
    n2 = nx[n2];
    ring[n1] += ring[n2];
    return (double)ring[n1] * (1.0 / (double)ULONG_MAX);
  }
  double gauss() {
    const double two_pi = 6.28318530717959;  // MY_2pi, MyRand.h:39
    if (!gen) {
      t = two_pi * ranf();
      r = std::sqrt(-std::log(1.0 - ranf()));
      gen = true;
      return r * std::cos(t);
    }
    gen = false;
    return r * std::sin(t);
  }
};

// su(3) noise from eight uniforms (MyMath.cc:295-322); su(2) from four
// uniforms of which three are used (the reference reads a fourth element
// past the end of g[3] at MyMath.cc:352-361; here it is drawn explicitly).
void noise_from_uniforms(int nc, double* g, double* out /*2*nc*nc*/) {
  const double twopi = std::atan(1.0) * 8.0;
  const int ng = (nc == 3) ? 8 : 4;
  for (int i = 0; i < ng; i += 2) {
    const double t = twopi * g[i];
    const double r = std::sqrt(-std::log(1.0 - g[i + 1]));
    g[i] = r * std::cos(t);
    g[i + 1] = r * std::sin(t);
  }
  if (nc == 3) {
    g[7] *= std::sqrt(1.0 / 3.0);
    const double m[18] = {0, g[7] + g[6], g[1], g[0], g[3], g[2], -g[1], g[0], 0, g[7] - g[6],
                          g[5], g[4], -g[3], g[2], -g[5], g[4], 0, -g[7] * 2};
    std::memcpy(out, m, sizeof(m));
  } else {
    const double m[8] = {0, g[2], g[1], g[0], -g[1], g[0], 0, -g[2]};
    std::memcpy(out, m, sizeof(m));
  }
}

// su(3)/su(2) noise from RAND55 gaussians (MyMath.cc:256-292, 326-346).
void noise_rand55(int nc, Rand55& R, double* out) {
  cx P[9];
  for (auto& z : P) z = cx{0, 0};
  if (nc == 3) {
    double g = R.gauss();
    P[1].im = g; P[3].im = g;
    g = R.gauss();
    P[1].re = g; P[3].re = -g;
    g = R.gauss();
    P[2].im = g; P[6].im = g;
    g = R.gauss();
    P[2].re = g; P[6].re = -g;
    g = R.gauss();
    P[5].im = g; P[7].im = g;
    g = R.gauss();
    P[5].re = g; P[7].re = -g;
    g = R.gauss();
    const double gg = R.gauss() * std::sqrt(1.0 / 3.0);
    P[0].im = gg + g;
    P[4].im = gg - g;
    P[8].im = -2. * gg;
  } else {
    double g = R.gauss();
    P[1].im = g; P[2].im = g;
    g = R.gauss();
    P[1].re = g; P[2].re = -g;
    g = R.gauss();
    P[0].im = g; P[3].im = -g;
  }
  for (int e = 0; e < nc * nc; e++) {
    out[2 * e] = (double)P[e].re;
    out[2 * e + 1] = (double)P[e].im;
  }
}

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).
inline void philox4x32(uint32_t ctr[4], uint32_t key[2]) {
  uint32_t k0 = key[0], k1 = key[1];
  for (int round = 0; round < 10; round++) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * ctr[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * ctr[2];
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    const uint32_t c0 = hi1 ^ ctr[1] ^ k0, c2 = hi0 ^ ctr[3] ^ k1;
    ctr[0] = c0;
    ctr[1] = lo1;
    ctr[2] = c2;
    ctr[3] = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

// Device-noise convention: link l of step `step` uses counters
// (l, step_lo, step_hi, q) for q = 0,1 and key (seed_lo, seed_hi); each 32-bit
// word u maps to the uniform (u + 0.5) * 2^-32 in (0,1).
void noise_philox_link(int nc, uint64_t seed, uint64_t step, uint32_t link, double* out) {
  double g[8];
  const int nq = (nc == 3) ? 2 : 1;
  for (int q = 0; q < nq; q++) {
    uint32_t ctr[4] = {link, (uint32_t)step, (uint32_t)(step >> 32), (uint32_t)q};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    philox4x32(ctr, key);
    for (int w = 0; w < 4; w++) g[4 * q + w] = ((double)ctr[w] + 0.5) * (1.0 / 4294967296.0);
  }
  noise_from_uniforms(nc, g, out);
}

#endif  // !ORC_EXTENDED

// ---------------------------------------------------------------- planted mistakes
// Mutation switches for the tests that pin this oracle: a test that claims
// to pin a formula shows that it fails on an oracle carrying a plausible
// mistake in it.  Off (0, 1.0) everywhere except inside those tests.
//   fa:  1 alpha with the wrong sign, 2 the 1/V of the inverse DFT dropped,
//        3 p^2 = 4 sin^2(p) instead of 4 sin^2(p/2) (lattice.h:1144-1168)
//   rk2: factor on the nc*eps/6 coefficient of the RK2 corrector
struct Plant {
  int fa = 0;
  double rk2 = 1.0;
};
Plant g_plant;

// ---------------------------------------------------------------- step
struct Params {
  double eps;        // Langevin step epsilon (tau_g of Quench.cfg)
  double alpha;      // gauge-fixing alpha (reference sign conventions)
  int integrator;    // 0 Euler, 1 RK2
  int update;        // 0 Jacobi, 1 on-the-fly (sequential, in place),
                     // 2 on the fly in checkerboard x direction order
  int gauge_fix;     // 0 none, 1 Fourier accelerated, 2 default in-place,
                     // 3 default in-place in checkerboard site order
  int zero_modes;    // 1: zero_modes_subtraction
  double c1 = 0.0;   // rectangle coefficient: 0 Wilson, -1/12 TLSYM, -0.331 IWA, -1.4088 DBW2
};

#if !ORC_EXTENDED
// ---------------------------------------------------------------- adjoint
// Generalised series with a matrix at order 0 (random configurations with a
// Haar-random order 0): orders k = 0..N, each NC x NC, [(k*NE + e)].  The
// Wilson action truncated at order K is f = sum_{x, mu<nu} Re Tr P_{mu nu}(x)_K
// with P = U_mu(x) U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag (the plaquettes of
// nspt.cc:1327-1346).  Its gradient is taken in reverse mode, each
// plaquette's products recomputed in the backward pass.  Complex gradient
// convention: G = df/dRe U + i df/dIm U, so df = Re Tr(G^dag dU).
struct GSer {
  int nc, N;
  std::vector<cx> v;  // (N+1) * nc*nc
  GSer(int nc_, int N_) : nc(nc_), N(N_), v((size_t)(N_ + 1) * nc_ * nc_, cx{0, 0}) {}
  int ne() const { return nc * nc; }
  cx* ord(int k) { return v.data() + (size_t)k * ne(); }
  const cx* ord(int k) const { return v.data() + (size_t)k * ne(); }
};

// full Cauchy product C_k = sum_{i+j=k} A_i B_j, k = 0..N
GSer gmul(const GSer& A, const GSer& B) {
  GSer C(A.nc, A.N);
  for (int i = 0; i <= A.N; i++)
    for (int j = 0; i + j <= A.N; j++) mat_mac(A.nc, A.ord(i), B.ord(j), C.ord(i + j));
  return C;
}
GSer gdag(const GSer& A) {
  GSer D(A.nc, A.N);
  for (int k = 0; k <= A.N; k++) mat_dag(A.nc, A.ord(k), D.ord(k));
  return D;
}
// adjoints of C = A B: Abar_i = sum_k Cbar_k B_{k-i}^dag, Bbar_j = sum_k A_{k-j}^dag Cbar_k
GSer gadj_a(const GSer& Cbar, const GSer& B) {
  const GSer Bd = gdag(B);
  GSer Ab(B.nc, B.N);
  for (int i = 0; i <= B.N; i++)
    for (int k = i; k <= B.N; k++) mat_mac(B.nc, Cbar.ord(k), Bd.ord(k - i), Ab.ord(i));
  return Ab;
}
GSer gadj_b(const GSer& A, const GSer& Cbar) {
  const GSer Ad = gdag(A);
  GSer Bb(A.nc, A.N);
  for (int j = 0; j <= A.N; j++)
    for (int k = j; k <= A.N; k++) mat_mac(A.nc, Ad.ord(k - j), Cbar.ord(k), Bb.ord(j));
  return Bb;
}
void gacc(GSer& a, const GSer& b) {
  for (size_t i = 0; i < a.v.size(); i++) a.v[i] += b.v[i];
}

struct GField {
  int nc, N;
  const Latt* lat;
  double* data;  // [(4*site + mu) * (N+1)*NE] complex
  size_t S() const { return (size_t)(N + 1) * nc * nc; }
  GSer get(int site, int mu) const {
    GSer g(nc, N);
    std::memcpy(g.v.data(), data + 2 * (4 * (size_t)site + mu) * S(), sizeof(cx) * S());
    return g;
  }
};

double re_tr(int nc, const cx* m) {
  double t = 0;
  for (int q = 0; q < nc; q++) t += m[q * nc + q].re;
  return t;
}

// Re Tr P_{mu nu}(x)_K
double gplaq(const GField& U, int x, int mu, int nu, int K) {
  const Latt& Z = *U.lat;
  const GSer P = gmul(gmul(gmul(U.get(x, mu), U.get(Z.up[4 * x + mu], nu)), gdag(U.get(Z.up[4 * x + nu], mu))),
                      gdag(U.get(x, nu)));
  return re_tr(U.nc, P.ord(K));
}

// reverse-mode gradient of f_K into G (same layout as U)
void gaction_grad(const GField& U, int K, double* Gout) {
  const Latt& Z = *U.lat;
  const int nc = U.nc, N = U.N;
  std::vector<GSer> G(4 * (size_t)Z.V, GSer(nc, N));
  GSer Pbar(nc, N);
  for (int q = 0; q < nc; q++) Pbar.ord(K)[q * nc + q] = cx{1, 0};
  for (int x = 0; x < Z.V; x++)
    for (int mu = 0; mu < 4; mu++)
      for (int nu = mu + 1; nu < 4; nu++) {
        const int l1 = 4 * x + mu, l2 = 4 * Z.up[4 * x + mu] + nu, l3 = 4 * Z.up[4 * x + nu] + mu, l4 = 4 * x + nu;
        const GSer U1 = U.get(x, mu), U2 = U.get(l2 / 4, nu), U3d = gdag(U.get(l3 / 4, mu)),
                   U4d = gdag(U.get(x, nu));
        // forward: T1 = U1 U2, T2 = T1 U3^dag, P = T2 U4^dag
        const GSer T1 = gmul(U1, U2);
        const GSer T2 = gmul(T1, U3d);
        // backward
        const GSer T2b = gadj_a(Pbar, U4d);
        gacc(G[l4], gdag(gadj_b(T2, Pbar)));
        const GSer T1b = gadj_a(T2b, U3d);
        gacc(G[l3], gdag(gadj_b(T1, T2b)));
        gacc(G[l1], gadj_a(T1b, U2));
        gacc(G[l2], gadj_b(U1, T1b));
      }
  for (size_t l = 0; l < G.size(); l++) std::memcpy(Gout + 2 * l * U.S(), G[l].v.data(), sizeof(cx) * U.S());
}

// sum of the staples of link (x, mu) (QCDenvNODEpt_test.h:554-557), matrix order 0
GSer gstaple_sum(const GField& U, int x, int mu) {
  const Latt& Z = *U.lat;
  GSer S(U.nc, U.N);
  for (int nu = 0; nu < 4; nu++) {
    if (nu == mu) continue;
    const int xpmu = Z.up[4 * x + mu], xpnu = Z.up[4 * x + nu], xmnu = Z.dn[4 * x + nu];
    const int xmnu_pmu = Z.up[4 * xmnu + mu];
    gacc(S, gmul(gmul(U.get(xpmu, nu), gdag(U.get(xpnu, mu))), gdag(U.get(x, nu))));
    gacc(S, gmul(gmul(gdag(U.get(xmnu_pmu, nu)), gdag(U.get(xmnu, mu))), U.get(xmnu, nu)));
  }
  return S;
}

#endif  // !ORC_EXTENDED

struct Geom {
  int L[4], nc, N;
};

void zero_modes(Field& F, std::vector<Ser>& MM) {
  const Latt& Z = *F.lat;
  for (int mu = 0; mu < 4; mu++) ser_scale(MM[mu], real(1) / Z.V);  // MM *= rVol
#pragma omp parallel for schedule(dynamic, 64) if (g_par)
  for (long l = 0; l < 4L * Z.V; l++) {
    const int s = (int)(l / 4), mu = (int)(l % 4);
    Ser W = ser_sub(ser_log(F.get(s, mu)), MM[mu]);
    ser_reH(W);
    F.set(s, mu, ser_exp(W));
  }
}

void fa_gauge_fixing(Field& F, double alpha) {
  const Latt& Z = *F.lat;
  const int S = F.S, nc = F.nc, N = F.N;
  std::vector<cx> Wg((size_t)Z.V * S, cx{0, 0});
#pragma omp parallel for schedule(dynamic, 16) if (g_par)
  for (int x = 0; x < Z.V; x++) {
    Ser W(nc, N);
    for (int mu = 0; mu < 4; mu++) {
      Ser d = ser_sub(ser_log(F.get(x, mu)), ser_log(F.get(Z.dn[x * 4 + mu], mu)));
      for (size_t q = 0; q < W.v.size(); q++) W.v[q] += d.v[q];
    }
    std::memcpy(&Wg[(size_t)x * S], W.v.data(), sizeof(cx) * S);
  }
  // plan howmany = nc*nc*PTORD + 1, stride = S, dist = 1 (Quenched.cc:827-838)
  dft_many(Z.L, S, S, Wg.data(), -1);
  for (int q = 0; q < S; q++) Wg[q] = 0.0 * Wg[q];  // PBC: k = 0 mode removed
  std::vector<real> p2[4];
  for (int d = 0; d < 4; d++) {
    p2[d].resize(Z.L[d]);
    for (int i = 0; i < Z.L[d]; i++) {
      const real p = 2 * kPi * i / Z.L[d];
      real h = std::sin(g_plant.fa == 3 ? p : p / 2);
      h *= 4 * h;
      p2[d][i] = h;  // lattice.h:1144-1168
    }
  }
#pragma omp parallel for schedule(static) if (g_par)
  for (int x = 1; x < Z.V; x++) {
    int c[4];
    Z.coords(x, c);
    const real a = g_plant.fa == 1 ? -alpha : alpha;
    const real f = -(a / (p2[0][c[0]] + p2[1][c[1]] + p2[2][c[2]] + p2[3][c[3]]));
    for (int q = 0; q < S; q++) Wg[(size_t)x * S + q] = f * Wg[(size_t)x * S + q];
  }
  dft_many(Z.L, S, S, Wg.data(), +1);
  std::vector<Ser> G(Z.V);
#pragma omp parallel for schedule(dynamic, 16) if (g_par)
  for (int x = 0; x < Z.V; x++) {
    Ser W(nc, N);
    std::memcpy(W.v.data(), &Wg[(size_t)x * S], sizeof(cx) * S);
    if (g_plant.fa != 2) ser_div(W, (real)Z.V);
    ser_reH(W);
    G[x] = ser_exp(W);
  }
#pragma omp parallel for schedule(dynamic, 64) if (g_par)
  for (long l = 0; l < 4L * Z.V; l++) {
    const int x = (int)(l / 4), mu = (int)(l % 4);
    F.set(x, mu, ser_mul(ser_mul(G[x], F.get(x, mu)), ser_dag(G[Z.up[x * 4 + mu]])));
  }
}

// Default (non-FA) stochastic gauge fixing: sequential and in place
// (nspt.cc:731-805).
void sgf_default(Field& F, double alpha) {
  const Latt& Z = *F.lat;
  for (int s = 0; s < Z.V; s++) {
    Ser W(F.nc, F.N);
    for (int mu = 0; mu < 4; mu++) {
      Ser d = ser_sub(F.get(s, mu), F.get(Z.bwd(s, mu), mu));
      for (size_t q = 0; q < W.v.size(); q++) W.v[q] += d.v[q];
    }
    W = ser_sub(W, ser_dag(W));
    ser_trless(W);
    Ser Wa = W, Wm = W;
    ser_scale(Wa, alpha);
    ser_scale(Wm, -alpha);
    Ser G1 = ser_exp(Wa), G2 = ser_exp(Wm);
    for (int mu = 0; mu < 4; mu++) {
      F.set(s, mu, ser_mul(G1, F.get(s, mu)));
      const int b = Z.bwd(s, mu);
      F.set(b, mu, ser_mul(F.get(b, mu), G2));
    }
  }
}

// parity of a site, (x0+x1+x2+x3) mod 2
int site_parity(const Latt& Z, int s) {
  int x[4];
  Z.coords(s, x);
  return (x[0] + x[1] + x[2] + x[3]) & 1;
}

// The sequential gauge fixing of nspt.cc:660-862 with its sites taken in
// checkerboard order: all sites of parity 0, then of parity 1.  Sites of one
// parity share no link, so each pass is order-independent (the GPU form).
void sgf_checkerboard(Field& F, double alpha) {
  const Latt& Z = *F.lat;
  for (int par = 0; par < 2; par++)
    for (int s = 0; s < Z.V; s++) {
      if (site_parity(Z, s) != par) continue;
      Ser W(F.nc, F.N);
      for (int mu = 0; mu < 4; mu++) {
        Ser d = ser_sub(F.get(s, mu), F.get(Z.bwd(s, mu), mu));
        for (size_t q = 0; q < W.v.size(); q++) W.v[q] += d.v[q];
      }
      W = ser_sub(W, ser_dag(W));
      ser_trless(W);
      Ser Wa = W, Wm = W;
      ser_scale(Wa, alpha);
      ser_scale(Wm, -alpha);
      Ser G1 = ser_exp(Wa), G2 = ser_exp(Wm);
      for (int mu = 0; mu < 4; mu++) {
        F.set(s, mu, ser_mul(G1, F.get(s, mu)));
        const int b = Z.bwd(s, mu);
        F.set(b, mu, ser_mul(F.get(b, mu), G2));
      }
    }
}

// RK2 corrector drift: (D1 + D2)/2 times (1 + (nc eps/6) x^2), x = beta^-1/2.
Ser rk2_combine(const Ser& D1, const Ser& D2, int nc, double eps) {
  Ser D = ser_add(D1, D2);
  ser_scale(D, 0.5);
  const real c = (real)nc * eps / 6 * g_plant.rk2;
  for (int k = D.N; k >= 3; k--) {
    cx* o = D.ord(k);
    const cx* p = D.ord(k - 2);
    for (int e = 0; e < D.ne(); e++) o[e] += c * p[e];
  }
  return D;
}

// One NsptEvolve step (nspt.cc:963-1033).  eta: 4V noise matrices in link
// order.  obs: plaq[2(N+1)], norm[N], trU[8N] (may be null).
void step(const Geom& g, real* Udata, const Params& p, const double* eta, double* obs_out) {
  Latt Z(g.L);
  const int S = 1 + g.nc * g.nc * g.N, ne2 = 2 * g.nc * g.nc;
  Field U{g.nc, g.N, S, &Z, Udata};
  const real stau = std::sqrt((real)p.eps), tau_g = -(real)p.eps / (2 * g.nc);
  std::vector<Ser> MM(4, Ser(g.nc, g.N));
  Obs obs;
  obs.init(g.N, p.c1 != 0.0);

  if (p.update == 2) {
    // on the fly in checkerboard order: for parity 0 then 1, for mu = 0..3,
    // every link (x, mu) of that parity at once (their Wilson staples use no
    // link of the same direction and parity)
    for (int par = 0; par < 2; par++)
      for (int mu = 0; mu < 4; mu++)
        for (int s = 0; s < Z.V; s++) {
          if (site_parity(Z, s) != par) continue;
          Ser W = drift(U, s, mu, tau_g, &obs, p.c1);
          sub_noise(W, stau, eta + (size_t)(4 * s + mu) * ne2);
          Ser Un = ser_mul(ser_exp(W), U.get(s, mu));
          U.set(s, mu, Un);
          MM[mu] = ser_add(MM[mu], ser_log(Un));
        }
  } else if (p.update == 1) {
    // on the fly: in place, lexicographic site order, mu inner (nspt.cc:365-441)
    for (int s = 0; s < Z.V; s++)
      for (int mu = 0; mu < 4; mu++) {
        Ser W = drift(U, s, mu, tau_g, &obs, p.c1);
        sub_noise(W, stau, eta + (size_t)(4 * s + mu) * ne2);
        Ser Un = ser_mul(ser_exp(W), U.get(s, mu));
        U.set(s, mu, Un);
        MM[mu] = ser_add(MM[mu], ser_log(Un));
      }
  } else {
    // Jacobi: every link from the frozen field, links in parallel; the
    // observable and zero-mode sums are then formed in link order
    const long nl = 4L * Z.V;
    std::vector<real> Fbuf((size_t)8 * Z.V * S);
    Field Fn{g.nc, g.N, S, &Z, Fbuf.data()};
    std::vector<Ser> D1, lg(nl);
    std::vector<Obs::Rec> rec(obs_out ? nl : 0);
    if (p.integrator == 1) {
      // predictor U1 = exp(D(U) - sqrt(eps) eta) U
      D1.resize(nl);
#pragma omp parallel for schedule(dynamic, 16) if (g_par)
      for (long l = 0; l < nl; l++) {
        const int s = (int)(l / 4), mu = (int)(l % 4);
        D1[l] = drift(U, s, mu, tau_g, nullptr, p.c1, obs_out ? &rec[l] : nullptr);
        Ser W = D1[l];
        sub_noise(W, stau, eta + (size_t)l * ne2);
        Fn.set(s, mu, ser_mul(ser_exp(W), U.get(s, mu)));
      }
    }
    std::vector<real> Gbuf(p.integrator == 1 ? (size_t)8 * Z.V * S : 0);
#pragma omp parallel for schedule(dynamic, 16) if (g_par)
    for (long l = 0; l < nl; l++) {
      const int s = (int)(l / 4), mu = (int)(l % 4);
      Ser W;
      if (p.integrator == 1) {
        // corrector: drift at U1, combined with D(U); applied to U
        W = rk2_combine(D1[l], drift(Fn, s, mu, tau_g, nullptr, p.c1), g.nc, p.eps);
      } else {
        W = drift(U, s, mu, tau_g, nullptr, p.c1, obs_out ? &rec[l] : nullptr);
      }
      sub_noise(W, stau, eta + (size_t)l * ne2);
      Ser Un = ser_mul(ser_exp(W), U.get(s, mu));
      if (p.integrator == 1)
        std::memcpy(Gbuf.data() + 2 * (size_t)l * S, Un.v.data(), sizeof(cx) * S);
      else
        Fn.set(s, mu, Un);
      lg[l] = ser_log(Un);
    }
    for (long l = 0; l < nl; l++) {
      if (obs_out) obs.add(rec[l]);
      MM[l % 4] = ser_add(MM[l % 4], lg[l]);
    }
    // Umu = Fmu (nspt.cc:992)
    if (p.integrator == 1)
      std::memcpy(Udata, Gbuf.data(), sizeof(real) * Gbuf.size());
    else
      std::memcpy(Udata, Fbuf.data(), sizeof(real) * Fbuf.size());
  }
  if (p.zero_modes) zero_modes(U, MM);
  if (p.gauge_fix == 1) fa_gauge_fixing(U, p.alpha);
  if (p.gauge_fix == 2) sgf_default(U, p.alpha);
  if (p.gauge_fix == 3) sgf_checkerboard(U, p.alpha);
  if (obs_out) obs.write(g.nc, Z.V, obs_out);
}

#if !ORC_EXTENDED
// plaquette_measure, OpenMP form (nspt.cc:1355-1402): sum over mu != nu of
// Tr(U_mu * staple) / (72 V) (SU(3)) or (48 V) (SU(2)).
void plaquette_measure(const Geom& g, const double* Udata, double* out) {
  Latt Z(g.L);
  const int S = 1 + g.nc * g.nc * g.N;
  Field U{g.nc, g.N, S, &Z, const_cast<double*>(Udata)};
  std::vector<cx> w(g.N + 1, cx{0, 0}), t(g.N + 1);
  for (int s = 0; s < Z.V; s++)
    for (int mu = 0; mu < 4; mu++) {
      Ser W = ser_mul(U.get(s, mu), staple_sum(U, s, mu));
      ser_tr(W, t.data());
      for (int k = 0; k <= g.N; k++) w[k] += t[k];
    }
  const double nrm = (g.nc == 3) ? 72.0 * Z.V : 48.0 * Z.V;
  for (int k = 0; k <= g.N; k++) {
    out[2 * k] = w[k].re / nrm;
    out[2 * k + 1] = w[k].im / nrm;
  }
}

// __K_MOM_ANALYSIS__ momentum norms (nspt.cc:919-944): the forward DFT of
// every component of the field, as planGluon[0] (Quenched.cc:792-815: the
// 4 (1 + nc^2 N) complex of a site, FFTW_FORWARD, unnormalised), then per
// order k = 1..N and momentum p (site order)
//   out[(k-1) V + p] = sum_mu Re Tr(Ut_mu,k(p) Ut_mu,k(p)^dag),
// the trace as the sum over rows of Re (A A^dag)_rr.  The field is read only
// (the reference transforms back and multiplies by 1/V, nspt.cc:936-941).
void knorm(const Geom& g, const double* Udata, double* out) {
  Latt Z(g.L);
  const int S = 1 + g.nc * g.nc * g.N, NE = g.nc * g.nc;
  std::vector<cx> T((size_t)Z.V * 4 * S);
  for (size_t i = 0; i < T.size(); i++) T[i] = cx{Udata[2 * i], Udata[2 * i + 1]};
  dft_many(g.L, 4 * S, 4 * S, T.data(), -1);
  for (int k = 0; k < g.N; k++)
    for (int p = 0; p < Z.V; p++) {
      double acc = 0.0;
      for (int mu = 0; mu < 4; mu++) {
        const cx* m = T.data() + (size_t)(4 * p + mu) * S + 1 + (size_t)k * NE;
        double tr = 0.0;
        for (int r = 0; r < g.nc; r++) {
          double d = 0.0;
          for (int c = 0; c < g.nc; c++) {
            const cx v = m[r * g.nc + c];
            d += v.re * v.re + v.im * v.im;  // Re (v conj(v))
          }
          tr += d;
        }
        acc += tr;
      }
      out[(size_t)k * Z.V + p] = acc;
    }
}

// ComputeLoops (nspt.cc:1156-1312): per plane mu < nu and site x, the sums
//   w21 += 2x1 + 1x2 rectangles, w22 += 2x2, w42 += 4x2 + 2x4, w44 += 4x4,
// with the reference's neighbour chain (a .. t) and product order; out gets
// Tr of each sum (QCDpt.cc:115-122), 4 x (N+1) complex, unnormalised.
void compute_loops(const Geom& gm, const double* Udata, double* out) {
  Latt Z(gm.L);
  const int nc = gm.nc, N = gm.N;
  const int S = 1 + nc * nc * N;
  Field F{nc, N, S, &Z, const_cast<double*>(Udata)};
  auto U = [&](int site, int mu) { return F.get(site, mu); };
  auto up = [&](int site, int mu) { return Z.fwd(site, mu); };
  Ser w21(nc, N), w22(nc, N), w42(nc, N), w44(nc, N);
  for (int mu = 0; mu < 3; mu++)
    for (int nu = mu + 1; nu < 4; nu++)
      for (int x = 0; x < Z.V; x++) {
        const int a = up(x, mu), b = up(a, mu), c = up(b, nu), e = up(x, nu), d = up(e, mu);
        const int f = up(c, nu), g = up(d, nu), h = up(e, nu);
        const Ser bl = ser_mul(U(x, mu), U(a, mu));
        const Ser ml = ser_mul(U(h, mu), U(g, mu));
        const Ser cd = ser_mul(U(b, nu), U(c, nu));
        const Ser ld = ser_mul(U(x, nu), U(e, nu));
        w22 = ser_add(w22, ser_mul(ser_mul(bl, cd), ser_dag(ser_mul(ld, ml))));
        const Ser r21 = ser_mul(ser_mul(bl, U(b, nu)), ser_dag(ser_mul(ser_mul(U(x, nu), U(e, mu)), U(d, mu))));
        const Ser r12 =
            ser_mul(ser_mul(ser_mul(U(x, mu), U(a, nu)), U(d, nu)), ser_dag(ser_mul(ld, U(h, mu))));
        w21 = ser_add(w21, ser_add(r21, r12));
        const int i = up(h, nu), j = up(f, nu), k = up(j, nu), n = up(i, nu), l = up(n, mu);
        const int m = up(b, mu), o = up(m, mu), p = up(o, nu), q = up(p, nu), r = up(f, mu);
        const int s = up(q, nu), t = up(k, mu);
        const Ser br = ser_mul(U(b, mu), U(m, mu));
        const Ser mr = ser_mul(U(f, mu), U(r, mu));
        const Ser tr = ser_mul(U(k, mu), U(t, mu));
        const Ser tl = ser_mul(U(n, mu), U(l, mu));
        const Ser lu = ser_mul(U(h, nu), U(i, nu));
        const Ser cu = ser_mul(U(f, nu), U(j, nu));
        const Ser ru = ser_mul(U(q, nu), U(s, nu));
        const Ser rd = ser_mul(U(o, nu), U(p, nu));
        const Ser in42 = ser_add(ser_mul(ser_mul(br, rd), ser_dag(ser_mul(ml, mr))),
                                 ser_mul(ser_mul(cd, cu), ser_dag(ser_mul(lu, tl))));
        w42 = ser_add(w42, ser_mul(ser_mul(bl, in42), ser_dag(ld)));
        const Ser top = ser_mul(ser_mul(ser_mul(bl, br), rd), ru);
        const Ser bot = ser_mul(ser_mul(ser_mul(ld, lu), tl), tr);
        w44 = ser_add(w44, ser_mul(top, ser_dag(bot)));
      }
  const Ser* ws[4] = {&w21, &w22, &w42, &w44};
  std::vector<cx> tt(N + 1);
  for (int w = 0; w < 4; w++) {
    ser_tr(*ws[w], tt.data());
    for (int kk = 0; kk <= N; kk++) {
      out[2 * (w * (N + 1) + kk)] = tt[kk].re;
      out[2 * (w * (N + 1) + kk) + 1] = tt[kk].im;
    }
  }
}

Ser from_ptr(int nc, int N, const double* p) {
  Ser s(nc, N);
  std::memcpy(s.v.data(), p, sizeof(cx) * s.v.size());
  return s;
}
void to_ptr(const Ser& s, double* p) { std::memcpy(p, s.v.data(), sizeof(cx) * s.v.size()); }

#endif  // !ORC_EXTENDED

Geom geom_from(const int* gi) {
  Geom g;
  for (int d = 0; d < 4; d++) g.L[d] = gi[d];
  g.nc = gi[4];
  g.N = gi[5];
  return g;
}

#if !ORC_EXTENDED
// Long run for statistics: nsteps steps of p from U (in place), the noise of
// each step drawn from one ranlxd stream (level 1) in link order.  bins
// receives, per bin of nsteps / nbin consecutive steps, the mean over the
// bin of the real part of every plaquette order 0..N (the observable of
// each step, measured on the field entering it, nspt.cc:392-402, 570-597).
void run_stats(const Geom& g, double* U, const Params& p, int seed, long nsteps, int nbin, double* bins) {
  Ranlxd R(seed, 1);
  const int V = g.L[0] * g.L[1] * g.L[2] * g.L[3];
  const int ne2 = 2 * g.nc * g.nc, ng = (g.nc == 3) ? 8 : 4;
  const int olen = 2 * (g.N + 1) + g.N + 8 * g.N + (p.c1 != 0.0 ? 2 * (g.N + 1) : 0);
  std::vector<double> eta((size_t)4 * V * ne2), obs(olen);
  const long per = nsteps / nbin;
  for (int b = 0; b < nbin; b++) {
    double* acc = bins + (size_t)b * (g.N + 1);
    for (int k = 0; k <= g.N; k++) acc[k] = 0.0;
    for (long s = 0; s < per; s++) {
      for (int l = 0; l < 4 * V; l++) {
        double u[8];
        for (int i = 0; i < ng; i++) u[i] = R.next();
        noise_from_uniforms(g.nc, u, eta.data() + (size_t)l * ne2);
      }
      step(g, U, p, eta.data(), obs.data());
      for (int k = 0; k <= g.N; k++) acc[k] += obs[2 * k];
    }
    for (int k = 0; k <= g.N; k++) acc[k] /= (double)per;
  }
}
#endif  // !ORC_EXTENDED

}  // namespace

#if !ORC_EXTENDED
// ================================================================ C ABI
// geometry arrays are int[6] = {L0, L1, L2, L3, nc, N}; params arrays are
// double[7] = {eps, alpha, integrator, update, gauge_fix, zero_modes, c1}.
extern "C" {

// planted mistakes (see Plant): fa variant, factor on the RK2 coefficient
void orc_set_plant(int fa, double rk2) {
  g_plant.fa = fa;
  g_plant.rk2 = rk2;
}

void orc_run_stats(const int* gi, double* U, const double* pp, int seed, long nsteps, int nbin, double* bins) {
  Params p{pp[0], pp[1], (int)pp[2], (int)pp[3], (int)pp[4], (int)pp[5], pp[6]};
  run_stats(geom_from(gi), U, p, seed, nsteps, nbin, bins);
}

void orc_ser_mul(int nc, int N, const double* a, const double* b, double* c) {
  to_ptr(ser_mul(from_ptr(nc, N, a), from_ptr(nc, N, b)), c);
}
void orc_ser_exp(int nc, int N, const double* a, double* c) { to_ptr(ser_exp(from_ptr(nc, N, a)), c); }
// Seconds for `rounds` rounds of n_mul products a*b, n_exp exp(a) and n_log
// log(b) (the speed calibration against the reference's own algebra,
// oracle/calibrate.py); the results feed a checksum.
double orc_algebra_seconds(int nc, int N, const double* a, const double* b, int rounds, int n_mul, int n_exp,
                           int n_log, double* checksum) {
  const Ser A = from_ptr(nc, N, a), B = from_ptr(nc, N, b);
  double sum = 0.0;
  const auto t0 = std::chrono::steady_clock::now();
  for (int r = 0; r < rounds; r++) {
    for (int i = 0; i < n_mul; i++) sum += ser_mul(A, B).v[1].re;
    for (int i = 0; i < n_exp; i++) sum += ser_exp(A).v[1].re;
    for (int i = 0; i < n_log; i++) sum += ser_log(B).v[1].re;
  }
  const auto t1 = std::chrono::steady_clock::now();
  *checksum = sum;
  return std::chrono::duration<double>(t1 - t0).count();
}
void orc_ser_log(int nc, int N, const double* a, double* c) { to_ptr(ser_log(from_ptr(nc, N, a)), c); }
void orc_ser_reh(int nc, int N, const double* a, double* c) {
  Ser s = from_ptr(nc, N, a);
  ser_reH(s);
  to_ptr(s, c);
}
void orc_ser_trless(int nc, int N, const double* a, double* c) {
  Ser s = from_ptr(nc, N, a);
  ser_trless(s);
  to_ptr(s, c);
}
void orc_ser_dag(int nc, int N, const double* a, double* c) { to_ptr(ser_dag(from_ptr(nc, N, a)), c); }
void orc_ser_tr(int nc, int N, const double* a, double* tt) {
  ser_tr(from_ptr(nc, N, a), reinterpret_cast<cx*>(tt));
}

void orc_staple_sum(const int* gi, const double* U, int site, int mu, double* out) {
  Geom g = geom_from(gi);
  Latt Z(g.L);
  Field F{g.nc, g.N, 1 + g.nc * g.nc * g.N, &Z, const_cast<double*>(U)};
  to_ptr(staple_sum(F, site, mu), out);
}

// D[4V*S] = tau_g * reH(U * staples) for every link (no noise).
void orc_staple2x1(const int* gi, const double* U, int site, int mu, int nu, double* out) {
  Geom g = geom_from(gi);
  Latt Z(g.L);
  const int S = 1 + g.nc * g.nc * g.N;
  Field F{g.nc, g.N, S, &Z, const_cast<double*>(U)};
  to_ptr(staple2x1(F, site, mu, nu), out);
}
void orc_drift_c1(const int* gi, const double* U, double eps, double c1, double* D) {
  Geom g = geom_from(gi);
  Latt Z(g.L);
  const int S = 1 + g.nc * g.nc * g.N;
  Field F{g.nc, g.N, S, &Z, const_cast<double*>(U)};
  for (int s = 0; s < Z.V; s++)
    for (int mu = 0; mu < 4; mu++)
      to_ptr(drift(F, s, mu, -eps / (2.0 * g.nc), nullptr, c1), D + 2 * (size_t)(4 * s + mu) * S);
}
void orc_drift(const int* gi, const double* U, double eps, double* D) {
  Geom g = geom_from(gi);
  Latt Z(g.L);
  const int S = 1 + g.nc * g.nc * g.N;
  Field F{g.nc, g.N, S, &Z, const_cast<double*>(U)};
  for (int s = 0; s < Z.V; s++)
    for (int mu = 0; mu < 4; mu++)
      to_ptr(drift(F, s, mu, -eps / (2.0 * g.nc), nullptr), D + 2 * (size_t)(4 * s + mu) * S);
}

void orc_step(const int* gi, double* U, const double* pp, const double* eta, double* obs) {
  Params p{pp[0], pp[1], (int)pp[2], (int)pp[3], (int)pp[4], (int)pp[5], pp[6]};
  step(geom_from(gi), U, p, eta, obs);
}

// Stage entry points (Jacobi, Euler): gauge_wilson into F with MM[4*S] and
// observables; zero modes given MM; gauge fixings in place.
void orc_gauge_wilson_jacobi(const int* gi, const double* U, double* F, const double* eta, double eps,
                             double* MMout, double* obs) {
  Geom g = geom_from(gi);
  Latt Z(g.L);
  const int S = 1 + g.nc * g.nc * g.N, ne2 = 2 * g.nc * g.nc;
  Field Uf{g.nc, g.N, S, &Z, const_cast<double*>(U)};
  Field Ff{g.nc, g.N, S, &Z, F};
  std::vector<Ser> MM(4, Ser(g.nc, g.N));
  Obs o;
  o.init(g.N);
  for (int s = 0; s < Z.V; s++)
    for (int mu = 0; mu < 4; mu++) {
      Ser W = drift(Uf, s, mu, -eps / (2.0 * g.nc), &o);
      sub_noise(W, std::sqrt(eps), eta + (size_t)(4 * s + mu) * ne2);
      Ser Un = ser_mul(ser_exp(W), Uf.get(s, mu));
      Ff.set(s, mu, Un);
      MM[mu] = ser_add(MM[mu], ser_log(Un));
    }
  if (MMout)
    for (int mu = 0; mu < 4; mu++) to_ptr(MM[mu], MMout + 2 * (size_t)mu * S);
  if (obs) o.write(g.nc, Z.V, obs);
}

void orc_zero_modes(const int* gi, double* U, const double* MMin) {
  Geom g = geom_from(gi);
  Latt Z(g.L);
  const int S = 1 + g.nc * g.nc * g.N;
  Field F{g.nc, g.N, S, &Z, U};
  std::vector<Ser> MM(4);
  for (int mu = 0; mu < 4; mu++) MM[mu] = from_ptr(g.nc, g.N, MMin + 2 * (size_t)mu * S);
  zero_modes(F, MM);
}

void orc_fa_gauge_fixing(const int* gi, double* U, double alpha) {
  Geom g = geom_from(gi);
  Latt Z(g.L);
  Field F{g.nc, g.N, 1 + g.nc * g.nc * g.N, &Z, U};
  fa_gauge_fixing(F, alpha);
}

void orc_sgf_default(const int* gi, double* U, double alpha) {
  Geom g = geom_from(gi);
  Latt Z(g.L);
  Field F{g.nc, g.N, 1 + g.nc * g.nc * g.N, &Z, U};
  sgf_default(F, alpha);
}

void orc_compute_loops(const int* gi, const double* U, double* out) { compute_loops(geom_from(gi), U, out); }
void orc_sgf_checkerboard(const int* gi, double* U, double alpha) {
  Geom g = geom_from(gi);
  Latt Z(g.L);
  const int S = 1 + g.nc * g.nc * g.N;
  Field F{g.nc, g.N, S, &Z, U};
  sgf_checkerboard(F, alpha);
}
// Twisted boundary conditions for the calls that follow: mask bit d twists
// direction d with the nc x nc matrix om + 2 nc^2 d (row-major, interleaved
// complex); mask 0 restores periodic conditions.  The FA gauge fixing and
// the adjoint functions stay periodic.
void orc_set_twist(int nc, int mask, const double* om) {
  g_tw.mask = mask;
  g_tw.nc = nc;
  for (int d = 0; d < 4; d++)
    for (int q = 0; q < nc * nc; q++) {
      const bool on = (mask >> d) & 1;
      g_tw.om[d][q] = on ? cx{om[2 * (d * nc * nc + q)], om[2 * (d * nc * nc + q) + 1]}
                         : cx{(q / nc == q % nc) ? 1.0 : 0.0, 0.0};
    }
}
void orc_plaquette_measure(const int* gi, const double* U, double* out) {
  plaquette_measure(geom_from(gi), U, out);
}
void orc_knorm(const int* gi, const double* U, double* out) { knorm(geom_from(gi), U, out); }

// data: V * stride complex; transforms components 0..howmany-1.
void orc_dft_many(const int* L, int howmany, int stride, double* data, int sign) {
  dft_many(L, howmany, stride, reinterpret_cast<cx*>(data), sign);
}

// ---- RNG
void* orc_ranlxd_new(int seed, int level) { return new Ranlxd(seed, level); }
void orc_ranlxd_free(void* r) { delete static_cast<Ranlxd*>(r); }
void orc_ranlxd_draw(void* r, int n, double* out) {
  Ranlxd* g = static_cast<Ranlxd*>(r);
  for (int i = 0; i < n; i++) out[i] = g->next();
}
void* orc_rand55_new(long seed) { return new Rand55(seed); }
void orc_rand55_free(void* r) { delete static_cast<Rand55*>(r); }
void orc_rand55_ranf(void* r, int n, double* out) {
  for (int i = 0; i < n; i++) out[i] = static_cast<Rand55*>(r)->ranf();
}
void orc_rand55_gauss(void* r, int n, double* out) {
  for (int i = 0; i < n; i++) out[i] = static_cast<Rand55*>(r)->gauss();
}
// nlinks noise matrices (2*nc*nc doubles each) in consumption order.
void orc_noise_ranlxd(void* r, int nc, int nlinks, double* out) {
  Ranlxd* g = static_cast<Ranlxd*>(r);
  const int ng = (nc == 3) ? 8 : 4;
  for (int l = 0; l < nlinks; l++) {
    double u[8];
    for (int i = 0; i < ng; i++) u[i] = g->next();
    noise_from_uniforms(nc, u, out + (size_t)l * 2 * nc * nc);
  }
}
void orc_noise_rand55(void* r, int nc, int nlinks, double* out) {
  for (int l = 0; l < nlinks; l++) noise_rand55(nc, *static_cast<Rand55*>(r), out + (size_t)l * 2 * nc * nc);
}
void orc_noise_philox(int nc, unsigned long long seed, unsigned long long step, int nlinks, double* out) {
  for (int l = 0; l < nlinks; l++) noise_philox_link(nc, seed, step, (uint32_t)l, out + (size_t)l * 2 * nc * nc);
}
// ranlxd seed of the reference driver: first glibc rand() after srand(seed)
// (Quenched.cc:301-306).
int orc_glibc_first_rand(unsigned int seed) {
  srand(seed);
  return rand();
}

// ---- adjoint (generalised series, matrix order 0)
void orc_gser_mul(int nc, int N, const double* a, const double* b, double* out) {
  GSer A(nc, N), B(nc, N);
  std::memcpy(A.v.data(), a, sizeof(cx) * A.v.size());
  std::memcpy(B.v.data(), b, sizeof(cx) * B.v.size());
  const GSer C = gmul(A, B);
  std::memcpy(out, C.v.data(), sizeof(cx) * C.v.size());
}
// f_K = sum_{x, mu<nu} Re Tr P_K (site < 0), or the sum over the six
// plaquettes that contain link (site, mu) (the part of f_K that depends on it)
double orc_gaction(const int* gi, const double* U, int K, int site, int mu) {
  const Geom g = geom_from(gi);
  Latt Z(g.L);
  GField F{g.nc, g.N, &Z, const_cast<double*>(U)};
  double f = 0;
  if (site < 0) {
    for (int x = 0; x < Z.V; x++)
      for (int m = 0; m < 4; m++)
        for (int n = m + 1; n < 4; n++) f += gplaq(F, x, m, n, K);
    return f;
  }
  for (int nu = 0; nu < 4; nu++) {
    if (nu == mu) continue;
    const int a = std::min(mu, nu), b = std::max(mu, nu);
    f += gplaq(F, site, a, b, K);
    f += gplaq(F, Z.dn[4 * site + nu], a, b, K);
  }
  return f;
}
void orc_gaction_grad(const int* gi, const double* U, int K, double* G) {
  const Geom g = geom_from(gi);
  Latt Z(g.L);
  GField F{g.nc, g.N, &Z, const_cast<double*>(U)};
  gaction_grad(F, K, G);
}
void orc_gstaple_sum(const int* gi, const double* U, int site, int mu, double* out) {
  const Geom g = geom_from(gi);
  Latt Z(g.L);
  GField F{g.nc, g.N, &Z, const_cast<double*>(U)};
  const GSer S = gstaple_sum(F, site, mu);
  std::memcpy(out, S.v.data(), sizeof(cx) * S.v.size());
}

}  // extern "C"
#else
// ================================================================ extended C ABI
// The step in long double, as the accuracy reference of FP64 steps: the field
// (4V series, link order) comes and goes as a double-double pair, value =
// hi + lo, so nothing of the extended result is lost on return; the noise and
// the parameters are the double inputs a FP64 step reads.  obs may be null.
extern "C" {
void orcx_step(const int* gi, double* Uhi, double* Ulo, const double* pp, const double* eta, double* obs) {
  const Geom g = geom_from(gi);
  const size_t n = (size_t)8 * g.L[0] * g.L[1] * g.L[2] * g.L[3] * (1 + g.nc * g.nc * g.N);
  std::vector<real> U(n);
  for (size_t i = 0; i < n; i++) U[i] = (real)Uhi[i] + (real)Ulo[i];
  Params p{pp[0], pp[1], (int)pp[2], (int)pp[3], (int)pp[4], (int)pp[5], pp[6]};
  step(g, U.data(), p, eta, obs);
  for (size_t i = 0; i < n; i++) {
    Uhi[i] = (double)U[i];
    Ulo[i] = (double)(U[i] - (real)Uhi[i]);
  }
}
}  // extern "C"
#endif  // ORC_EXTENDED
