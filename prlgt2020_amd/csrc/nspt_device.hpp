// nspt_device.hpp -- device-side building blocks of the NSPT Langevin step
// for gfx950 (CDNA4), FP64 VALU, no MFMA.
//
// Lane mapping.  One wavefront (64 lanes) is one workgroup.  A link series
// (NC x NC complex matrices for orders 1..N plus a scalar order-0 flag) is
// spread over NE = NC*NC lanes: lane (j, e) owns matrix element e = NC*r + c
// of link slot j for *all* orders, in registers.  SU(3): 7 links x 9 lanes
// (lane 63 idle); SU(2): 16 links x 4 lanes.
//
// Truncated product C = A*B (the Cauchy product of QCDpt.h:92-114): the
// left operand A is staged in LDS (each lane writes its own element), the
// right operand B stays in HBM.  Lane (r, c) walks the orders of B in
// windows of QB: it loads column c of B (or of B^dag) for the window into
// registers, then streams rows r of A (or A^dag) from LDS and accumulates
// C_{t+q} += A_t[r][:] . B_q[:][c].  Every order index is a compile-time
// constant; the truncation bound is a wave-uniform runtime value, so one
// compiled instance serves every runtime order up to N.
//
// exp and log (QCDpt.h:176-220) are evaluated in Horner form with the same
// product; the Horner polynomial commutes with its generator, so
//   exp:  S <- 1 + (S*X)/k                 k = n..1
//   log:  S <- c_i + S*X,  c_i = (-1)^{i+1}/i, then log = S*X
// each step truncated at the order it can still reach.  The flop count is
// the reference's (N-1)N(N+1)/6 matrix products per exp/log.
//
// HBM layout of a series field with P series (structure of arrays):
//   data[(k*NE + e)*P + idx]  double2, k = 0..N-1 is order k+1
//   flag[idx]                 double2
#pragma once
#include <hip/hip_runtime.h>

namespace nspt {

struct SerField {
  double2* data;
  double2* flag;  // may be null for fields whose flag is implied
  long P;         // number of series (plane stride)
};

__device__ __forceinline__ double2 c_add(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 c_sub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 c_mul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.y * b.x + a.x * b.y);
}
__device__ __forceinline__ double2 c_scale(double s, double2 a) { return make_double2(s * a.x, s * a.y); }
__device__ __forceinline__ double2 c_conj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ void c_mac(double2& acc, double2 a, double2 b) {
  acc.x = fma(a.x, b.x, acc.x);
  acc.x = fma(-a.y, b.y, acc.x);
  acc.y = fma(a.x, b.y, acc.y);
  acc.y = fma(a.y, b.x, acc.y);
}
__device__ __forceinline__ double2 zero2() { return make_double2(0.0, 0.0); }
__device__ __forceinline__ double2 one2() { return make_double2(1.0, 0.0); }

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Order window of the right operand (complex registers: 3*QB + N + 3).
template <int N>
struct QWin {
  static constexpr int value = N <= 8 ? N : (N <= 16 ? 8 : 5);
};

// Right operand of a product: series idx of a field in HBM, optionally
// daggered, optionally with -stau*eta subtracted at order 1 (the Langevin
// noise of nspt.cc:420-424, eta[idx*NE + e]).
struct BOp {
  const double2* data;
  long P;
  long idx;
  bool dag;
  const double2* eta;
  double stau;
};

__device__ __forceinline__ BOp bop(const SerField& f, long idx, bool dag = false) {
  return BOp{f.data, f.P, idx, dag, nullptr, 0.0};
}

// Lane bookkeeping for one wave = one workgroup.
template <int NC, int N>
struct Lane {
  static constexpr int NE = NC * NC;
  static constexpr int LPB = 64 / NE;  // link slots per wave
  static constexpr int SLOT = N * NE;  // LDS complex per slot
  int j, e, r, c;
  bool lds_ok;  // owns an LDS cell (false only for SU(3) lane 63)
  double2* slot;

  __device__ __forceinline__ Lane(double2* lds) {
    const int lane = threadIdx.x;
    lds_ok = lane < LPB * NE;
    j = lds_ok ? lane / NE : 0;
    e = lds_ok ? lane % NE : 0;
    r = e / NC;
    c = e % NC;
    slot = lds + j * SLOT;
  }

  // LDS staging of a lane-owned series (all orders)
  __device__ __forceinline__ void stage(const double2 (&X)[N]) const {
    if (lds_ok) {
#pragma unroll
      for (int k = 0; k < N; k++) slot[k * NE + e] = X[k];
    }
  }

  // Element (r,c) of the staged series or of its dagger, order index k.
  __device__ __forceinline__ double2 elem(bool dag, int k) const {
    if (!dag) return slot[k * NE + e];
    return c_conj(slot[k * NE + c * NC + r]);
  }
};

template <int NC>
__device__ __forceinline__ double2 pick(const double2 (&v)[NC], int i) {
  double x = v[0].x, y = v[0].y;
#pragma unroll
  for (int m = 1; m < NC; m++) {
    const double vx = v[m].x, vy = v[m].y;
    x = (i == m) ? vx : x;
    y = (i == m) ? vy : y;
  }
  return make_double2(x, y);
}

// Load of element `off` (32-bit, per lane) relative to a wave-uniform base;
// written so that the compiler can use scalar-base addressing.
__device__ __forceinline__ double2 ld_u(const double2* base, unsigned off) {
  return *reinterpret_cast<const double2*>(reinterpret_cast<const char*>(base) + (size_t)(off * 16u));
}
__device__ __forceinline__ void st_u(double2* base, unsigned off, double2 v) {
  *reinterpret_cast<double2*>(reinterpret_cast<char*>(base) + (size_t)(off * 16u)) = v;
}

// C = A*B truncated at order M+1 (0-based output index <= M), A staged in
// LDS with flag a0 (rows of A^dag if dagA), B in HBM with flag b0.  Adds the
// scalar-flag terms a0*B_k + b0*A_k (QCDpt.h:104-112) and returns a0*b0.
// ok: this lane's link is valid (loads of invalid links read zero).
// M must be wave-uniform and < N.
template <int NC, int N>
__device__ __forceinline__ double2 mul(const Lane<NC, N>& ln, bool dagA, const BOp& B, double2 a0, double2 b0,
                                       double2 (&C)[N], int M, int nrt, bool ok) {
  constexpr int NE = NC * NC;
  constexpr int QB = QWin<N>::value;
  M = uniform(M);
  int oa[NC];
  unsigned ob[NC];
#pragma unroll
  for (int m = 0; m < NC; m++) {
    oa[m] = dagA ? (m * NC + ln.r) : (ln.r * NC + m);
    const int e = B.dag ? (ln.c * NC + m) : (m * NC + ln.c);
    ob[m] = (unsigned)(e * B.P + B.idx);
  }
  const double sa = dagA ? -1.0 : 1.0;
  const double sb = B.dag ? -1.0 : 1.0;
  const long plane = (long)NE * B.P;
#pragma unroll
  for (int k = 0; k < N; k++) C[k] = zero2();

#pragma unroll
  for (int q0 = 0; q0 < N; q0 += QB) {
    if (q0 <= M) {
      double2 Bw[QB][NC];
#pragma unroll
      for (int qq = 0; qq < QB; qq++) {
        const int q = q0 + qq;  // order q+1
        const double2* bq = B.data + q * plane;
#pragma unroll
        for (int m = 0; m < NC; m++) {
          double2 v = zero2();
          if (q < N) {
            if (ok && q < nrt) v = ld_u(bq, ob[m]);
            v.y *= sb;
            if (q == 0 && B.eta && ok) {
              const int e = B.dag ? (ln.c * NC + m) : (m * NC + ln.c);
              v = c_sub(v, c_scale(B.stau, B.eta[B.idx * NE + e]));
            }
          }
          Bw[qq][m] = v;
        }
      }
      // a0 * B_k
#pragma unroll
      for (int qq = 0; qq < QB; qq++)
        if (q0 + qq < N && q0 + qq <= M) C[q0 + qq] = c_add(C[q0 + qq], c_mul(a0, pick<NC>(Bw[qq], ln.r)));
      // Cauchy terms: A order t (LDS index t-1) times B order q0+qq+1
#pragma unroll
      for (int t = 1; t < N; t++) {
        if (t + q0 < N && t + q0 <= M) {
          const double2* row = ln.slot + (t - 1) * NE;
          double2 a[NC];
#pragma unroll
          for (int m = 0; m < NC; m++) {
            a[m] = row[oa[m]];
            a[m].y *= sa;
          }
#pragma unroll
          for (int qq = 0; qq < QB; qq++) {
            const int o = t + q0 + qq;
            if (o < N) {
              if (o <= M) {
#pragma unroll
                for (int m = 0; m < NC; m++) c_mac(C[o], a[m], Bw[qq][m]);
              }
            }
          }
        }
      }
    }
  }
  // b0 * A_k
#pragma unroll
  for (int k = 0; k < N; k++)
    if (k <= M) C[k] = c_add(C[k], c_mul(b0, ln.elem(dagA, k)));
  return c_mul(a0, b0);
}

// Horner evaluation of exp(X) (IS_EXP) or log(1 + X) for X = B (flag of B
// ignored, as exp and log of QCDpt.h:176-220 ignore it).  Result orders in S;
// the result flag is 1 (exp) or 0 (log).  Uses the LDS slot; the caller must
// have synchronised before the call; returns synchronised.
template <int NC, int N, bool IS_EXP>
__device__ __forceinline__ void horner(const Lane<NC, N>& ln, const BOp& B, double2 (&S)[N], int nrt, bool ok) {
  const int n = uniform(min(N, nrt));
#pragma unroll
  for (int k = 0; k < N; k++) S[k] = zero2();
  double2 flag = IS_EXP ? one2() : make_double2((n % 2 == 1 ? 1.0 : -1.0) / (double)n, 0.0);
  for (int s = 0; s < n; s++) {
    ln.stage(S);
    __syncthreads();
    mul<NC, N>(ln, false, B, flag, zero2(), S, s, nrt, ok);
    __syncthreads();
    if (IS_EXP) {
      const double a = 1.0 / (double)(n - s);
#pragma unroll
      for (int k = 0; k < N; k++) S[k] = c_scale(a, S[k]);
    } else {
      const int i = n - 1 - s;  // next coefficient c_i
      flag = (i >= 1) ? make_double2((i % 2 == 1 ? 1.0 : -1.0) / (double)i, 0.0) : zero2();
    }
  }
}

// reH of a staged series (QCDpt.cc:197-209): (M - M^dag)/2 minus Tr/NC.
template <int NC, int N>
__device__ __forceinline__ double2 reh_elem(const Lane<NC, N>& ln, int k) {
  constexpr int NE = NC * NC;
  const double2* m = ln.slot + k * NE;
  double2 d[NC];
#pragma unroll
  for (int q = 0; q < NC; q++) d[q] = c_scale(0.5, c_sub(m[q * NC + q], c_conj(m[q * NC + q])));
  double2 tr = d[0];
#pragma unroll
  for (int q = 1; q < NC; q++) tr = c_add(tr, d[q]);
  tr = c_scale(NC == 3 ? (1.0 / 3.0) : 0.5, tr);
  double2 v = c_scale(0.5, c_sub(m[ln.e], c_conj(m[ln.c * NC + ln.r])));
  if (ln.r == ln.c) v = c_sub(v, tr);
  return v;
}

// Global loads/stores of a lane-owned series.  Orders >= nrt read as zero and
// are not stored.
template <int NC, int N>
__device__ __forceinline__ void gload(const SerField& f, long idx, int e, int nrt, bool ok, double2 (&X)[N]) {
  constexpr int NE = NC * NC;
  const unsigned off = (unsigned)(e * f.P + idx);
  const long plane = (long)NE * f.P;
#pragma unroll
  for (int k = 0; k < N; k++) X[k] = (ok && k < nrt) ? ld_u(f.data + k * plane, off) : zero2();
}
template <int NC, int N>
__device__ __forceinline__ void gstore(const SerField& f, long idx, int e, int nrt, bool ok, const double2 (&X)[N]) {
  constexpr int NE = NC * NC;
  if (!ok) return;
  const unsigned off = (unsigned)(e * f.P + idx);
  const long plane = (long)NE * f.P;
#pragma unroll
  for (int k = 0; k < N; k++)
    if (k < nrt) st_u(f.data + k * plane, off, X[k]);
}
// f += X (own element), or f = X when first.
template <int NC, int N>
__device__ __forceinline__ void gaccum(const SerField& f, long idx, int e, int nrt, bool ok, const double2 (&X)[N],
                                       bool first) {
  constexpr int NE = NC * NC;
  if (!ok) return;
  const unsigned off = (unsigned)(e * f.P + idx);
  const long plane = (long)NE * f.P;
#pragma unroll
  for (int k = 0; k < N; k++)
    if (k < nrt) {
      double2* p = f.data + k * plane;
      st_u(p, off, first ? X[k] : c_add(ld_u(p, off), X[k]));
    }
}
__device__ __forceinline__ double2 gflag(const SerField& f, long idx, bool ok) {
  return (ok && f.flag) ? f.flag[idx] : zero2();
}

}  // namespace nspt
