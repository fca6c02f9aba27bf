// nspt_device.hpp -- device-side building blocks of the NSPT Langevin step
// for gfx950 (CDNA4), FP64 VALU, no MFMA.
//
// Lane mapping.  One wavefront (64 lanes) is one workgroup.  A link series
// (NC x NC complex matrices for orders 1..N plus a scalar order-0 flag) is
// spread over NE = NC*NC lanes: lane (j, e) owns matrix element e = NC*r + c
// of link slot j for *all* orders, in registers.  SU(3): 7 links x 9 lanes
// (lane 63 idle); SU(2): 16 links x 4 lanes.
//
// Truncated product C = A*B (the Cauchy product of QCDpt.h:92-114): both
// operands are staged in LDS (each lane writes its own element of A into
// slot A and of B into slot B, all orders, loads from HBM issued together).
// Lane (r, c) then walks the orders of B in windows of QB: it reads column c
// of B (or of B^dag) for the window into registers and streams rows r of A
// (or A^dag), accumulating C_{t+q} += A_t[r][:] . B_q[:][c].  Every order
// index is a compile-time constant; the truncation bound is a wave-uniform
// runtime value (one uniform branch per row), so one compiled instance
// serves every runtime order up to N.
//
// exp and log (QCDpt.h:176-220) are evaluated in Horner form with the same
// product; the Horner polynomial commutes with its generator, so
//   exp:  S <- 1 + (S*X)/k                 k = n..1
//   log:  S <- c_i + S*X,  c_i = (-1)^{i+1}/i, then log = S*X
// each step truncated at the order it can still reach.  The flop count is
// the reference's (N-1)N(N+1)/6 matrix products per exp/log.  The generator
// X stays in slot B for all steps.
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

// Order window of the right operand: QB orders of column c are read from LDS
// into registers (plus the next window, prefetched) for each pass over A.
#ifndef NSPT_QB
#define NSPT_QB 3
#endif
template <int N>
struct QWin {
  static constexpr int value = N <= 6 ? N : NSPT_QB;
};

// Lane bookkeeping for one wave = one workgroup.  LDS holds two slots per
// link (A and B operand), each N orders x NE elements.
template <int NC, int N>
struct Lane {
  static constexpr int NE = NC * NC;
  static constexpr int LPB = 64 / NE;  // link slots per wave
  static constexpr int SLOT = N * NE;  // LDS complex per slot
  static constexpr int LDS = 2 * LPB * SLOT;
  int j, e, r, c;
  bool lds_ok;  // owns an LDS cell (false only for SU(3) lane 63)
  double2* slotA;
  double2* slotB;

  __device__ __forceinline__ Lane(double2* lds) {
    const int lane = threadIdx.x;
    lds_ok = lane < LPB * NE;
    j = lds_ok ? lane / NE : 0;
    e = lds_ok ? lane % NE : 0;
    r = e / NC;
    c = e % NC;
    slotA = lds + j * SLOT;
    slotB = lds + LPB * SLOT + j * SLOT;
  }

  // LDS staging of a lane-owned series (all orders)
  __device__ __forceinline__ void stage(double2* slot, const double2 (&X)[N]) const {
    if (lds_ok) {
#pragma unroll
      for (int k = 0; k < N; k++) slot[k * NE + e] = X[k];
    }
  }

  // Element (r,c) of the series staged in slot A or of its dagger.
  __device__ __forceinline__ double2 elemA(bool dag, int k) const {
    if (!dag) return slotA[k * NE + e];
    return c_conj(slotA[k * NE + c * NC + r]);
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

// Load of element `off` (32-bit, per lane) of an array whose base is
// wave-uniform: a raw buffer load with a scalar resource descriptor, so the
// lane carries only a 32-bit offset.  Offsets past 2 GiB are not supported
// (every per-plane offset here is far below that).
__device__ __forceinline__ const double2* uniform_ptr(const double2* p) {
  const unsigned long long a = reinterpret_cast<unsigned long long>(p);
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)a);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(a >> 32));
  return reinterpret_cast<const double2*>(((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc(const double2* base) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<double2*>(uniform_ptr(base)), (short)0, -1, 0x00020000);
}
// 16-byte load at element `off` (per lane) + `soff` bytes (wave-uniform).
__device__ __forceinline__ double2 ld_r(__amdgpu_buffer_rsrc_t r, unsigned off, unsigned soff) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, off * 16u, (int)soff, 0);
  double2 d;
  __builtin_memcpy(&d, &v, sizeof(d));
  return d;
}
__device__ __forceinline__ double2 ld_u(const double2* base, unsigned off) { return ld_r(rsrc(base), off, 0u); }
__device__ __forceinline__ void st_u(double2* base, unsigned off, double2 v) {
  *reinterpret_cast<double2*>(reinterpret_cast<char*>(base) + (size_t)(off * 16u)) = v;
}

// C = A*B with A in slot A (A^dag if dagA, flag a0) and B in slot B (B^dag
// if dagB, flag b0).  Adds the scalar-flag terms a0*B_k + b0*A_k
// (QCDpt.h:104-112) and returns a0*b0.  Output orders up to M+1 (0-based
// index <= M, wave-uniform) are exact; entries above are left undefined and
// are never read by a caller (truncated products only feed upward).
// A^dag is handled as conj(sum A[m][r] conj(B[m][c])) so that the streamed
// rows are used as loaded.  has_b0 = false skips the b0 term (b0 = 0).
template <int NC, int N>
__device__ __forceinline__ double2 mul(const Lane<NC, N>& ln, bool dagA, bool dagB, double2 a0, double2 b0,
                                       bool has_b0, double2 (&C)[N], int M) {
  constexpr int NE = NC * NC;
  constexpr int QB = QWin<N>::value;
  constexpr int NW = (N + QB - 1) / QB;
  M = uniform(M);
  int oa[NC], ob[NC];
#pragma unroll
  for (int m = 0; m < NC; m++) {
    oa[m] = dagA ? (m * NC + ln.r) : (ln.r * NC + m);
    ob[m] = dagB ? (ln.c * NC + m) : (m * NC + ln.c);
  }
  const double sb = (dagB != dagA) ? -1.0 : 1.0;
  const double2 a0e = dagA ? c_conj(a0) : a0;

  auto loadwin = [&](int q0, double2 (&W)[QB][NC]) {
#pragma unroll
    for (int qq = 0; qq < QB; qq++) {
      const int q = q0 + qq;  // order q+1
#pragma unroll
      for (int m = 0; m < NC; m++) {
        double2 v = zero2();
        if (q < N) {
          v = ln.slotB[q * NE + ob[m]];
          v.y *= sb;
        }
        W[qq][m] = v;
      }
    }
  };
  auto rowload = [&](int t, double2 (&a)[NC]) {
    const double2* row = ln.slotA + (t - 1) * NE;
#pragma unroll
    for (int m = 0; m < NC; m++) a[m] = row[oa[m]];
  };

#pragma unroll
  for (int k = 0; k < N; k++) C[k] = zero2();
  // ping-pong buffers (static indices): window w uses Bw[w & 1], row t uses A[t & 1]
  double2 Bw[2][QB][NC];
  loadwin(0, Bw[0]);

#pragma unroll
  for (int w = 0; w < NW; w++) {
    const int q0 = w * QB;
    const int cur = w & 1;
    if (q0 > M) break;
    // Rows are re-read from LDS in every window: without this compiler
    // barrier the identical loads of earlier windows are reused and every
    // row stays live in registers across the whole product.
    asm volatile("" ::: "memory");
    if (w + 1 < NW) {
      if (q0 + QB <= M) loadwin(q0 + QB, Bw[cur ^ 1]);
    }
    // a0 * B_k
#pragma unroll
    for (int qq = 0; qq < QB; qq++)
      if (q0 + qq < N) C[q0 + qq] = c_add(C[q0 + qq], c_mul(a0e, pick<NC>(Bw[cur][qq], ln.r)));
    // Cauchy terms: A order t (LDS index t-1) times B order q0+qq+1 -> index t+q0+qq
    const int tmax = N - 1 - q0;
    double2 A[2][NC];
    if (tmax >= 1) rowload(1, A[1]);
#pragma unroll
    for (int t = 1; t < N; t++) {
      if (t <= tmax) {
        if (t + q0 > M) break;
        if (t + 1 <= tmax) rowload(t + 1, A[(t + 1) & 1]);
        // Every window order is formed for a row that reaches the bound:
        // products past order M+1 are left undefined and never read, and a
        // row without inner branches keeps 2*QB independent FMA chains.
#pragma unroll
        for (int qq = 0; qq < QB; qq++)
          if (t + q0 + qq < N) {
#pragma unroll
            for (int m = 0; m < NC; m++) c_mac(C[t + q0 + qq], A[t & 1][m], Bw[cur][qq][m]);
          }
        __builtin_amdgcn_sched_barrier(0);
      }
    }
  }
  if (dagA) {
#pragma unroll
    for (int k = 0; k < N; k++) C[k].y = -C[k].y;
  }
  if (has_b0) {
#pragma unroll
    for (int k = 0; k < N; k++) C[k] = c_add(C[k], c_mul(b0, ln.elemA(dagA, k)));
  }
  return c_mul(a0, b0);
}

// Horner evaluation of exp(X) (IS_EXP) or log(1 + X) for the generator X
// staged in slot B (its flag ignored, as exp and log of QCDpt.h:176-220
// ignore it).  Result orders in S; the result flag is 1 (exp) or 0 (log).
// n = number of orders formed (<= N).  The caller has synchronised after
// staging X; returns synchronised.
template <int NC, int N, bool IS_EXP>
__device__ __forceinline__ void horner(const Lane<NC, N>& ln, double2 (&S)[N], int n) {
  n = uniform(n);
#pragma unroll
  for (int k = 0; k < N; k++) S[k] = zero2();
  double2 flag = IS_EXP ? one2() : make_double2((n % 2 == 1 ? 1.0 : -1.0) / (double)n, 0.0);
  for (int s = 0; s < n; s++) {
    ln.stage(ln.slotA, S);
    __syncthreads();
    mul<NC, N>(ln, false, false, flag, zero2(), false, S, s);
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

// reH of the series staged in slot A (QCDpt.cc:197-209): (M - M^dag)/2 minus Tr/NC.
template <int NC, int N>
__device__ __forceinline__ double2 reh_elem(const Lane<NC, N>& ln, int k) {
  constexpr int NE = NC * NC;
  const double2* m = ln.slotA + k * NE;
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
// are not stored.  Loads are unmasked (lanes of invalid links hold clamped,
// valid indices); stores are masked by ok.
template <int NC, int N>
__device__ __forceinline__ void gload(const SerField& f, long idx, int e, int nrt, bool ok, double2 (&X)[N]) {
  constexpr int NE = NC * NC;
  const unsigned off = (unsigned)(e * f.P + idx);
  const long plane = (long)NE * f.P;
  (void)ok;
  const __amdgpu_buffer_rsrc_t r = rsrc(f.data);
#pragma unroll
  for (int k = 0; k < N; k++) X[k] = (k < nrt) ? ld_r(r, off, (unsigned)(k * plane * 16)) : zero2();
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
