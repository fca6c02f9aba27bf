// ref_harness.cpp -- C entry points around the reference's own series algebra
// and random-number sources, compiled together with the reference files
// MyMath.cc, MyRand.cc, QCDpt.cc and ranlxd/ranlxd.cc where they lie (see
// oracle/Makefile).  The result, oracle/_ref/libprlgt_ref.so, is test
// infrastructure: tests/golden/make_golden.py uses it to write golden vectors
// that pin the oracle.  It is never shipped and never loaded by the product.
//
// The reference's storage width is its compile-time allocORD (choices.h:30,
// = 6); a series crosses this boundary as 1 + nc*nc*6 complex doubles, the
// memory image of ptSU3 / ptSU2 (QCDpt.h:47-48, 657-658).  The runtime order
// PTORD (QCDpt.cc:31) is set through ref_set_ptord.

#include <chrono>
#include <cstring>

#include "QCDpt.h"

namespace {
template <class T>
T load(const double* p) {
  T s;
  std::memcpy(static_cast<void*>(&s), p, sizeof(T));
  return s;
}
template <class T>
void store(const T& s, double* p) {
  std::memcpy(p, static_cast<const void*>(&s), sizeof(T));
}
}  // namespace

extern "C" {

int ref_alloc_ord(void) { return allocORD; }
void ref_set_ptord(int n) { PTORD = n; }

// ---- ptSU3
void ref_su3_mul(const double* a, const double* b, double* c) { store(load<ptSU3>(a) * load<ptSU3>(b), c); }
void ref_su3_add(const double* a, const double* b, double* c) { store(load<ptSU3>(a) + load<ptSU3>(b), c); }
void ref_su3_exp(const double* a, double* c) { store(exp(load<ptSU3>(a)), c); }
void ref_su3_log(const double* a, double* c) { store(log(load<ptSU3>(a)), c); }
void ref_su3_dag(const double* a, double* c) { store(dag(load<ptSU3>(a)), c); }
void ref_su3_reh(const double* a, double* c) {
  ptSU3 s = load<ptSU3>(a);
  s.reH();
  store(s, c);
}
void ref_su3_trless(const double* a, double* c) {
  ptSU3 s = load<ptSU3>(a);
  s.Trless();
  store(s, c);
}
void ref_su3_tr(const double* a, double* tt) {
  ptSU3 s = load<ptSU3>(a);
  Cplx t[allocORD + 1];
  s.Tr(t);
  std::memcpy(tt, t, sizeof(Cplx) * (PTORD + 1));
}

// ---- ptSU2
void ref_su2_mul(const double* a, const double* b, double* c) { store(load<ptSU2>(a) * load<ptSU2>(b), c); }
void ref_su2_exp(const double* a, double* c) { store(exp(load<ptSU2>(a)), c); }
void ref_su2_log(const double* a, double* c) { store(log(load<ptSU2>(a)), c); }
void ref_su2_dag(const double* a, double* c) { store(dag(load<ptSU2>(a)), c); }
void ref_su2_reh(const double* a, double* c) {
  ptSU2 s = load<ptSU2>(a);
  s.reH();
  store(s, c);
}
void ref_su2_tr(const double* a, double* tt) {
  ptSU2 s = load<ptSU2>(a);
  Cplx t[allocORD + 1];
  s.Tr(t);
  std::memcpy(tt, t, sizeof(Cplx) * (PTORD + 1));
}

// ---- random sources
void ref_ranlxd(int seed, int level, int n, double* out) {
  ranlxd::Rand r(seed, level);
  r.ranlxd(out, out + n);
}
// n su(3) noise matrices from one ranlxd stream (MyMath.cc:295-322)
void ref_su3rand_ranlxd(int seed, int n, double* out) {
  ranlxd::Rand r(seed);
  for (int i = 0; i < n; i++) {
    SU3 m = SU3rand(r);
    std::memcpy(out + 18 * i, m.whr, sizeof(m.whr));
  }
}
void ref_rand55_ranf(long seed, int n, double* out) {
  MyRand r;
  r.init(seed);
  r.Rand(out, n);
}
void ref_rand55_gauss(long seed, int n, double* out) {
  MyRand r;
  r.init(seed);
  for (int i = 0; i < n; i++) out[i] = r.generate_gauss();
}
void ref_su3rand_rand55(long seed, int n, double* out) {
  MyRand r;
  r.init(seed);
  for (int i = 0; i < n; i++) {
    SU3 m = SU3rand(r);
    std::memcpy(out + 18 * i, m.whr, sizeof(m.whr));
  }
}
void ref_su2rand_rand55(long seed, int n, double* out) {
  MyRand r;
  r.init(seed);
  for (int i = 0; i < n; i++) {
    SU2 m = SU2rand(r);
    std::memcpy(out + 8 * i, m.whr, sizeof(m.whr));
  }
}

// Seconds for `rounds` rounds of the reference's SU(3) series algebra at the
// current PTORD: per round n_mul products c = a*b, n_exp exponentials and
// n_log logarithms of the series at a and b (memory images as above).  For
// the speed calibration of the oracle (oracle/calibrate.py); the results
// feed a checksum so that no operation is optimised away.
double ref_su3_algebra_seconds(const double* a, const double* b, int rounds, int n_mul, int n_exp, int n_log,
                               double* checksum) {
  const ptSU3 A = load<ptSU3>(a), B = load<ptSU3>(b);
  ptSU3 c;
  double sum = 0.0;
  const auto t0 = std::chrono::steady_clock::now();
  for (int r = 0; r < rounds; r++) {
    for (int i = 0; i < n_mul; i++) {
      c = A * B;
      sum += c.ptU[0].whr[0].re;
    }
    for (int i = 0; i < n_exp; i++) {
      c = exp(A);
      sum += c.ptU[0].whr[0].re;
    }
    for (int i = 0; i < n_log; i++) {
      c = log(B);
      sum += c.ptU[0].whr[0].re;
    }
  }
  const auto t1 = std::chrono::steady_clock::now();
  *checksum = sum;
  return std::chrono::duration<double>(t1 - t0).count();
}

}  // extern "C"
