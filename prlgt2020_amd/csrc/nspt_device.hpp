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
// slot A and of B into slot B, all orders, loads from HBM issued together;
// daggers are applied while staging).  Lane (r, c) then reads row r of A
// and column c of B tile by tile (QT orders of each) and accumulates
// C_{a+b+1} += A_a[r][:] . B_b[:][c] along anti-diagonal bands of tiles,
// highest band first, holding only the 2QT-1 outputs a band can reach.
// Finished orders go to registers or straight back into an operand slot
// (in place), so chained products need no restaging.  Every order index is
// a compile-time constant; the truncation bound is a wave-uniform runtime
// value, so one compiled instance serves every runtime order up to N.
//
// exp and log (QCDpt.h:176-220) are evaluated in Horner form with the same
// product; the Horner polynomial commutes with its generator, so
//   exp:  S <- 1 + (S*X)/k                 k = n..1
//   log:  S <- c_i + S*X,  c_i = (-1)^{i+1}/i, then log = S*X
// each step truncated at the order it can still reach.  The flop count is
// the reference's (N-1)N(N+1)/6 matrix products per exp/log.  The generator
// X stays in slot B and S in slot A for all steps.
//
// HBM layout of a series field with P series (structure of arrays in
// tiles of TILE = 64 consecutive series):
//   data[((idx/TILE)*nrt*NE + k*NE + e)*TILE + idx%TILE]  double2, k = 0..nrt-1
//                                                         is order k+1
//   flag[idx]                                             double2
#pragma once
#include <hip/hip_runtime.h>

// A large field (32^4 at N = 32: 19 GB) lies beyond the 32-bit byte offset
// of one buffer descriptor: every kernel instance is also compiled with
// NSPT_BIG_FIELDS = 1, which addresses each lane's series with 64 bits and
// which the host picks for fields past 2 GiB (the 64-bit lane addresses
// cost registers and address arithmetic the common case should not pay).
// Tile edge of the product: pairs of operand orders are taken QT x QT at a
// time.
#ifndef NSPT_QT
#define NSPT_QT 4
#endif
// Occupancy shape of a build: links per wave (0: the choice of
// links_per_wave below) and the waves per SIMD the kernels are compiled for
// (their register budget: 512 / waves registers per lane).
#ifndef NSPT_LPW
#define NSPT_LPW 0
#endif
#ifndef NSPT_WAVES_PER_EU
#define NSPT_WAVES_PER_EU 1
#endif
// Every build (field size x tile edge x occupancy shape) lives in an inline
// namespace of its own, so that its kernels, their host handles and its
// helpers are distinct symbols: two builds of one template kernel under one
// name would be merged by the linker into one handle, and every build would
// launch the code registered first.
#ifndef NSPT_BIG_FIELDS
#define NSPT_BIG_FIELDS 0
#endif
#define NSPT_VARIANT_NAME2(f, q, l, w) f##_q##q##_l##l##_w##w
#define NSPT_VARIANT_NAME(f, q, l, w) NSPT_VARIANT_NAME2(f, q, l, w)
#if NSPT_BIG_FIELDS
#define NSPT_VARIANT NSPT_VARIANT_NAME(big_fields, NSPT_QT, NSPT_LPW, NSPT_WAVES_PER_EU)
#else
#define NSPT_VARIANT NSPT_VARIANT_NAME(std_fields, NSPT_QT, NSPT_LPW, NSPT_WAVES_PER_EU)
#endif

namespace nspt {

struct SerField {
  double2* data;
  double2* flag;  // may be null for fields whose flag is implied
  long P;         // number of series
};

// A series field holds its series in tiles of NSPT_TILE consecutive
// indices: the tile of series idx holds every order and element of its
// NSPT_TILE series, order-major, then element, then series, so that order k,
// element e of series idx is complex entry ser_index(idx, k*NE + e).  One
// link's series then spans one tile (184 KB at 20 orders of SU(3)) instead
// of nrt*NE planes P*16 bytes apart: its loads touch one or two pages where
// the plane layout touched a page per order and element (the GPU's L1
// address-translation cache then missed 9-10% of the product kernels'
// accesses), and the NSPT_TILE series of one (k, e) are still contiguous.
// The allocation is rounded up to whole tiles.
constexpr int NSPT_TILE = 64;
constexpr int NSPT_TILE_SHIFT = 6;
__host__ __device__ inline long ser_alloc(long P, int nrt, int NE) {
  return (P + NSPT_TILE - 1) / NSPT_TILE * NSPT_TILE * nrt * NE;
}
__host__ __device__ inline long ser_index(long idx, int ke, int nrt, int NE) {
  return (idx >> NSPT_TILE_SHIFT) * ((long)nrt * NE * NSPT_TILE) + (long)ke * NSPT_TILE + (idx & (NSPT_TILE - 1));
}

inline namespace NSPT_VARIANT {

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

// The drift's products are full-length (truncated at the top order only), so
// a larger tile pays there; the Horner steps of exp/log are mostly short
// products, where the partial bands of a larger tile waste more than its
// extra reuse saves.
#ifndef NSPT_QT_DRIFT
#define NSPT_QT_DRIFT (NSPT_QT == 4 ? 5 : NSPT_QT)
#endif
#define NSPT_QT_MAX (NSPT_QT_DRIFT > NSPT_QT ? NSPT_QT_DRIFT : NSPT_QT)
#ifndef NSPT_SLOT_PAD
#define NSPT_SLOT_PAD 0
#endif

// Lane bookkeeping for one wave = one workgroup.  LDS holds two slots per
// link (A and B operand), each N orders x NE elements, plus QT-2 orders of
// padding after the last slot: a tile reads up to QT-2 orders past the end
// of a slot (into the next slot or the padding), always for products above
// the truncation bound, whose results are never used.
// Link slots per wave.  The kernels run one wave per SIMD (register bound),
// so a CU holds at most four; when the slots of 64/NE links leave LDS for
// fewer waves, fewer links per wave can pay: the choice maximises waves per
// CU (LDS in 1 KiB granules, 160 KiB per CU, plus a flag per link) times
// working lanes.  SU(3): 7 links up to N = 24, 4 at N = 32; SU(2): 16, but
// 15 at N = 20 and 13 at N = 24.  A build with NSPT_LPW > 0 takes that many
// (the occupancy variants the batch autotuner times).
template <int NC, int N>
constexpr int links_per_wave() {
  constexpr int NE = NC * NC;
  static_assert(NSPT_LPW >= 0 && NSPT_LPW * NE <= 64, "links per wave fit one wave");
  if (NSPT_LPW > 0) return NSPT_LPW;
  int best = 1, best_score = 0;
  for (int l = 64 / NE; l >= 1; l--) {
    const int bytes = (2 * l * (N * NE + NSPT_SLOT_PAD) + (NSPT_QT_MAX - 2) * NE + l) * 16;
    const int waves = 163840 / ((bytes + 1023) / 1024 * 1024);
    const int score = (waves < 4 ? waves : 4) * l * NE;
    if (score > best_score) {
      best_score = score;
      best = l;
    }
  }
  return best;
}

template <int NC, int N>
struct Lane {
  static constexpr int NE = NC * NC;
  static constexpr int LPB = links_per_wave<NC, N>();
  static constexpr int SLOT = N * NE + NSPT_SLOT_PAD;  // LDS complex per slot (pitch)
  static constexpr int PAD = (NSPT_QT_MAX > 2 ? NSPT_QT_MAX - 2 : 0) * NE;
  static constexpr int LDS = 2 * LPB * SLOT + PAD;
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

  // LDS staging of a lane-owned series (all orders), or of its dagger: the
  // lane that owns (r, c) writes conj(X) at (c, r).  With rev, order k of
  // the slot receives order N-1-k of X.
  __device__ __forceinline__ void stage(double2* slot, const double2 (&X)[N], bool dag = false,
                                        bool rev = false) const {
    if (lds_ok) {
      const int pos = dag ? (c * NC + r) : e;
      const double sy = dag ? -1.0 : 1.0;
#pragma unroll
      for (int k = 0; k < N; k++) {
        const double2 v = rev ? X[N - 1 - k] : X[k];
        slot[k * NE + pos] = make_double2(v.x, sy * v.y);
      }
    }
  }
};

// Load of element `off` (32-bit, per lane) of an array whose base is
// wave-uniform: a raw buffer load with a scalar resource descriptor, so the
// lane carries only a 32-bit offset.  Offsets past 2 GiB are not supported
// (fields past 2 GiB take the NSPT_BIG_FIELDS build).
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
__device__ __forceinline__ void st_u(double2* base, unsigned off, double2 v) {
  *reinterpret_cast<double2*>(reinterpret_cast<char*>(base) + (size_t)(off * 16u)) = v;
}

// Truncated product A*B for A staged in slot A (flag a0) and B in slot B
// (flag b0): the Cauchy product of QCDpt.h:92-114 with the flag terms
// a0*B_k + b0*A_k.  Returns the flag a0*b0.  has_b0 = false skips the b0
// term (b0 = 0).  Daggers are taken when staging (Lane::stage).
//
// Destination (wave-uniform choice):
//   to_lds: order k is written, times `scale` and daggered if `dag`, into
//           LDS slot `dst` -- which may be slot A or slot B itself: the
//           product is formed in place (see below);
//   else:   C[k] += order k (registers).
// Output order indices up to M (wave-uniform) are exact; entries above are
// undefined and never read by a caller (truncated products only feed
// upward).
//
// Order index a of A times index b of B lands on index a+b+1.  The pairs
// (a, b) are visited in QT x QT tiles along bands a0 + b0 = s, from the
// highest band down to s = 0.  Band s reaches outputs s+1 .. s+2QT-1 and
// reads operand orders below s+QT only, so 2QT-1 accumulators suffice;
// after band s the outputs s+QT .. s+2QT-1 are final and every operand
// order >= s is dead, which is what lets a finished order overwrite the
// operand slots.  Bands are unrolled (output indices are constants); the
// tiles of a band form a runtime loop.  Lane (r, c) reads row r of A and
// column c of B from LDS.
//
// OFF = 0 forms the full Cauchy product C_{a+b} of series whose order 0 is a
// matrix stored at index 0 (no flags; a0, b0 unused): the same bands with
// outputs shifted down by one.  It reads up to QT-1 (not QT-2) orders past a
// slot, so its LDS needs NE more complex of padding.
//
// ADD (with to_lds, dag false): order k is written as C[k] + scale*(A*B)_k,
// C holding this lane's element of an addend series.
//
// mul_bands is the band loop alone, for any OFF: every output index k in
// [OFF, N) whose band runs is passed once to retire(k, v) with its Cauchy
// sum v (exact up to M), after the last read of every operand index >= the
// band's s; outputs below OFF are the caller's.  A caller whose operands
// start at a higher order stores them from that order on (entry i of A
// holding order i + pa, of B order i + pb) and takes OFF = pa + pb - 1, so
// that no tile is spent on the zero orders.
template <int NC, int N, int OFF, int Q, class Retire>
__device__ __forceinline__ void mul_bands(const Lane<NC, N>& ln, int M, Retire&& retire) {
  constexpr int NE = NC * NC;
  static_assert(Q <= NSPT_QT_MAX, "LDS padding covers tiles up to NSPT_QT_MAX");
  constexpr int NO = 2 * Q - 1;
  M = uniform(M);
  const double2* pa = ln.slotA + ln.r * NC;
  const double2* pb = ln.slotB + ln.c;
  double2 acc[NO];
#pragma unroll
  for (int o = 0; o < NO; o++) acc[o] = zero2();

  double2 TA[Q][NC], TB[Q][NC];
  auto load = [&](int a, int b) {
    const double2* ra = pa + a * NE;
    const double2* cb = pb + b * NE;
    // columns of B first, then the rows of A in the order they are used,
    // so that the first row's products wait for fewer reads
#pragma unroll
    for (int d = 0; d < Q; d++)
#pragma unroll
      for (int m = 0; m < NC; m++) TB[d][m] = cb[d * NE + m * NC];
#pragma unroll
    for (int d = 0; d < Q; d++)
#pragma unroll
      for (int m = 0; m < NC; m++) TA[d][m] = ra[d * NE + m];
  };
  auto row = [&](int da) {
#pragma unroll
    for (int m = 0; m < NC; m++)
#pragma unroll
      for (int db = 0; db < Q; db++) c_mac(acc[da + db], TA[da][m], TB[db][m]);
  };
  // Rows da <= rem of a tile (all of them when rem >= QT - 1): a row whose
  // pairs all land above the bound is skipped; the pairs of a kept row that
  // land above it only feed outputs that are never read.  The empty asm
  // keeps each guarded row a real (uniform) branch instead of a select.
  auto tile = [&](int rem) {
    row(0);
#pragma unroll
    for (int da = 1; da < Q; da++) {
      if (rem >= Q - 1) {
        row(da);
      } else if (da <= rem) {
        asm volatile("");
        row(da);
      }
    }
  };
  auto band = [&](int s, int rem) {
    for (int a = 0; a <= s; a += Q) {
      load(a, s - a);
      tile(rem);
    }
  };

  // band s reaches outputs s+OFF .. s+OFF+2QT-2
  constexpr int TOP = N - 1 - OFF;
  constexpr int STOP = TOP >= 0 ? (TOP / Q) * Q : 0;
  // (a guard, not a break: the band loop must unroll fully)
#pragma unroll
  for (int s = STOP; s >= 0; s -= Q) {
    if (s + OFF <= M) {
      const int rem = M - OFF - s;
      if (rem >= Q - 1)
        band(s, Q - 1);
      else
        band(s, rem);
#pragma unroll
      for (int d = 0; d < Q; d++)
        if (s + OFF + Q - 1 + d < N) retire(s + OFF + Q - 1 + d, acc[Q - 1 + d]);
      // carry outputs s+OFF .. s+OFF+QT-2 to the next band down
#pragma unroll
      for (int o = NO - 1; o >= 0; o--) acc[o] = (o >= Q) ? acc[o - Q] : zero2();
    }
  }
  // outputs OFF .. OFF+QT-2 (carried past band 0, or flag terms only)
#pragma unroll
  for (int d = 0; d + 1 < Q; d++)
    if (OFF + d < N) retire(OFF + d, acc[Q + d]);
}

template <int NC, int N, int OFF = 1, bool ADD = false, int Q = NSPT_QT>
__device__ __forceinline__ double2 mul(const Lane<NC, N>& ln, double2 a0, double2 b0, bool has_b0,
                                       double2 (&C)[N], int M, bool to_lds, double2* dst, bool dag,
                                       double scale) {
  constexpr int NE = NC * NC;
  const double2* own_a = ln.slotA + ln.e;
  const double2* own_b = ln.slotB + ln.e;
  const int pos = dag ? (ln.c * NC + ln.r) : ln.e;
  const double sx = scale, sy = dag ? -scale : scale;
  // Order k is final with value v (flag terms still to add).  The flag
  // terms read order k of both operands before it is overwritten.
  auto retire = [&](int k, double2 v) {
    if (OFF == 1) {
      v = c_add(v, c_mul(a0, own_b[k * NE]));
      if (has_b0) v = c_add(v, c_mul(b0, own_a[k * NE]));
    }
    if (to_lds) {
      if (ln.lds_ok) {
        if (ADD)
          dst[k * NE + pos] = make_double2(fma(sx, v.x, C[k].x), fma(sy, v.y, C[k].y));
        else
          dst[k * NE + pos] = make_double2(sx * v.x, sy * v.y);
      }
    } else {
      C[k] = c_add(C[k], v);
    }
  };
  mul_bands<NC, N, OFF, Q>(ln, M, retire);
  // the flag-only index 0 of the flagged form
  if (OFF == 1) retire(0, zero2());
  return OFF == 1 ? c_mul(a0, b0) : zero2();
}

// Product into LDS (in place allowed) / added to registers.
template <int NC, int N, int Q = NSPT_QT>
__device__ __forceinline__ double2 mul_lds(const Lane<NC, N>& ln, double2 a0, double2 b0, bool has_b0, int M,
                                           double2* dst, bool dag, double scale = 1.0) {
  double2 unused[N];
  return mul<NC, N, 1, false, Q>(ln, a0, b0, has_b0, unused, M, true, dst, dag, scale);
}

// exp(X)*B by Horner from the right, T <- B + (X*T)/k for k = n..1, which
// costs what exp(X) alone costs (each step is truncated at the order it can
// still reach) and saves the product with B.  X (flag 0) must be staged in
// slot A and B in slot B; bv holds this lane's element of every order of B,
// bflag its flag.  T is formed in place in slot B, which holds exp(X)*B
// (flag bflag) on return.
template <int NC, int N>
__device__ __forceinline__ void horner_exp_times(const Lane<NC, N>& ln, int n, double2 bflag, double2 (&bv)[N]) {
  n = uniform(n);
  for (int s = 0; s < n; s++) {
    __syncthreads();
    mul<NC, N, 1, true>(ln, zero2(), bflag, true, bv, s, true, ln.slotB, false, 1.0 / (double)(n - s));
  }
  __syncthreads();
}

// Horner evaluation of exp(X) (IS_EXP) or log(1 + X) for the generator X
// staged in slot B (its flag ignored, as exp and log of QCDpt.h:176-220
// ignore it).  The result orders are left in slot A; the result flag is 1
// (exp) or 0 (log).  n = number of orders formed (<= N).  Each step forms
// S*X in place in slot A.
template <int NC, int N, bool IS_EXP>
__device__ __forceinline__ void horner(const Lane<NC, N>& ln, int n) {
  constexpr int NE = NC * NC;
  n = uniform(n);
  if (ln.lds_ok) {
#pragma unroll
    for (int k = 0; k < N; k++) ln.slotA[k * NE + ln.e] = zero2();
  }
  double2 flag = IS_EXP ? one2() : make_double2((n % 2 == 1 ? 1.0 : -1.0) / (double)n, 0.0);
  for (int s = 0; s < n; s++) {
    __syncthreads();
    const double a = IS_EXP ? 1.0 / (double)(n - s) : 1.0;
    mul_lds<NC, N>(ln, flag, zero2(), false, s, ln.slotA, false, a);
    if (!IS_EXP) {
      const int i = n - 1 - s;  // next coefficient c_i
      flag = (i >= 1) ? make_double2((i % 2 == 1 ? 1.0 : -1.0) / (double)i, 0.0) : zero2();
    }
  }
  __syncthreads();
}

// ------------------------------------------------------------ exp / log by Paterson-Stockmeyer
// Instances of at least NSPT_PS_MIN_N orders evaluate exp and log as
// p(X) = sum_{i <= n} c_i X^i in the Paterson-Stockmeyer form (Paterson and
// Stockmeyer, SIAM J. Comput. 2 (1973) 60) with Y = X^3:
//   X^2 = X X,  Y = X^2 X,
//   R_J = q_J,  R_j = q_j + R_{j+1} Y  (j = J-1 .. 0, J = n/3),  p(X) = R_0,
//   q_j = c_{3j} + c_{3j+1} X + c_{3j+2} X^2,
// R_j truncated at order n - 3j (it is multiplied by Y j more times).  The
// q_j cost no product (each lane adds its own elements), and every product
// is stored from its operands' lowest order on (X^2 from order 2, Y from
// order 3), so no tile is spent on their zero orders: at n = 20, 739 pairs
// of orders against the 1330 of the Horner form (376 against 680 at 16,
// 175 against 286 at 12).  The same terms are summed in another association
// than the reference's Horner loop (QCDpt.h:176-220): as accurate (against
// an extended-precision evaluation, exp to 2e-16 and log within the Horner
// form's own error), and a step's agreement with the oracle is what it is
// with the Horner form (on evolved fields at 12 and 16 orders the worst
// per-order difference is the same to two digits: the step's conditioning
// sets it).  Below 12 orders the products are too short for the fewer
// pairs to pay for the extra product set-ups.
#ifndef NSPT_PS_MIN_N
#define NSPT_PS_MIN_N 12
#endif
template <int N>
constexpr bool use_ps() {
  return N >= NSPT_PS_MIN_N;
}
// tile edge of its products: they are longer than the Horner steps, and the
// larger tile pays (as in the drift)
#ifndef NSPT_QT_PS
#define NSPT_QT_PS NSPT_QT_DRIFT
#endif

// 1/i! with two roundings: i! is exact in double up to 22!, and i!/22! is
// an exact integer up to i = 32.
__host__ __device__ constexpr double inv_factorial(int i) {
  double f = 1.0, g = 1.0;
  for (int k = 2; k <= i; k++) {
    if (k <= 22)
      f *= k;
    else
      g *= k;
  }
  return (1.0 / f) / g;
}
struct InvFactorials {
  double v[33];
  constexpr InvFactorials() : v() {
    for (int i = 0; i < 33; i++) v[i] = inv_factorial(i);
  }
};
__constant__ constexpr InvFactorials kInvFact{};

// c_i of exp (1/i!) or of log(1 + X) ((-1)^{i+1}/i, c_0 = 0); 0 above n.
template <bool IS_EXP>
__device__ __forceinline__ double series_coef(int i, int n) {
  if (i > n) return 0.0;
  if (IS_EXP) return kInvFact.v[i];
  return i == 0 ? 0.0 : ((i & 1) ? 1.0 : -1.0) / (double)i;
}

// p(X) for the generator X (its flag ignored, as in horner()), given as this
// lane's element of every order; both slots are used.  On return slot A
// holds orders 1..n of exp(X) (IS_EXP, flag 1) or log(1 + X) (flag 0).
//
// after_stage() runs once X is staged, while the first product runs: a
// caller issues there the loads it needs only after the series.
struct NoHook {
  __device__ void operator()() const {}
};
template <int NC, int N, bool IS_EXP, class AfterStage = NoHook>
__device__ __forceinline__ void ps_series(const Lane<NC, N>& ln, int n, const double2 (&X)[N],
                                          AfterStage&& after_stage = NoHook{}) {
  constexpr int NE = NC * NC;
  constexpr int Q = NSPT_QT_PS;
  n = uniform(n);
  double2* const sa = ln.slotA;
  double2* const sb = ln.slotB;
  const int e = ln.e;
  const bool own = ln.lds_ok;
  // X^2 = X X into registers (order 1, index 0, is zero)
  double2 X2[N];
#pragma unroll
  for (int k = 0; k < N; k++) X2[k] = zero2();
  ln.stage(sa, X);
  ln.stage(sb, X);
  __syncthreads();
  after_stage();
  mul_bands<NC, N, 1, Q>(ln, n - 1, [&](int k, double2 v) { X2[k] = v; });
  __syncthreads();
  // X^2 from order 2 on into slot A; Y = X^2 X formed in place in slot B,
  // from order 3 on
  if (own) {
#pragma unroll
    for (int k = 1; k < N; k++) sa[(k - 1) * NE + e] = X2[k];
  }
  __syncthreads();
  mul_bands<NC, N, 2, Q>(ln, n - 1, [&](int k, double2 v) {
    if (own) sb[(k - 2) * NE + e] = v;
  });
  // Horner in Y, R_j formed in place in slot A (its flag r is c_{3j})
  const int J = n / 3;
  double r = series_coef<IS_EXP>(3 * J, n);
  {
    const double c1 = series_coef<IS_EXP>(3 * J + 1, n), c2 = series_coef<IS_EXP>(3 * J + 2, n);
    __syncthreads();
    if (own) {
#pragma unroll
      for (int k = 0; k < N; k++) sa[k * NE + e] = c_add(c_scale(c1, X[k]), c_scale(c2, X2[k]));
    }
  }
  for (int j = J - 1; j >= 0; j--) {
    __syncthreads();
    const double c1 = series_coef<IS_EXP>(3 * j + 1, n), c2 = series_coef<IS_EXP>(3 * j + 2, n), rj = r;
    // order k+1 of q_j + r_{j+1} Y, this lane's element
    auto addend = [&](int k) {
      double2 t = c_add(c_scale(c1, X[k]), c_scale(c2, X2[k]));
      if (k >= 2) t = c_add(t, c_scale(rj, sb[(k - 2) * NE + e]));
      return t;
    };
    mul_bands<NC, N, 3, Q>(ln, n - 3 * j - 1, [&](int k, double2 v) {
      if (own) sa[k * NE + e] = c_add(v, addend(k));
    });
    // orders 1..3 lie below the product's lowest order
    if (own) {
#pragma unroll
      for (int k = 0; k < 3 && k < N; k++) sa[k * NE + e] = addend(k);
    }
    r = series_coef<IS_EXP>(3 * j, n);
  }
  __syncthreads();
}

// exp(X)*B by ps_series and one product: X (flag 0) and B (flag bflag) as
// this lane's elements of every order.  On return slot B holds exp(X)*B
// (flag bflag), as after horner_exp_times().
// load_b() (if given) fills bv once X is staged; with dag_b the dagger of
// bv is staged as B.
template <int NC, int N, class LoadB = NoHook>
__device__ __forceinline__ void ps_exp_times(const Lane<NC, N>& ln, int n, const double2 (&X)[N], double2 bflag,
                                             const double2 (&bv)[N], LoadB&& load_b = NoHook{}, bool dag_b = false) {
  ps_series<NC, N, true>(ln, n, X, load_b);
  ln.stage(ln.slotB, bv, dag_b);
  __syncthreads();
  mul_lds<NC, N, NSPT_QT_PS>(ln, one2(), bflag, true, n - 1, ln.slotB, false);
  __syncthreads();
}

// exp(X)*B as ps_exp_times, with B staged into slot B by stage_b() once the
// series is formed: above 20 orders of SU(3) the registers of a held B (4N
// per lane) are what the series and its products need (holding it spilled
// 14-30 VGPRs at 24 orders and 112-132 at 32), so B is read again from
// memory instead.
template <int NC, int N, class StageB>
__device__ __forceinline__ void ps_exp_times_staged(const Lane<NC, N>& ln, int n, const double2 (&X)[N],
                                                    double2 bflag, StageB&& stage_b) {
  ps_series<NC, N, true>(ln, n, X);
  stage_b();
  __syncthreads();
  mul_lds<NC, N, NSPT_QT_PS>(ln, one2(), bflag, true, n - 1, ln.slotB, false);
  __syncthreads();
}
// Whether exp(X)*B holds B in registers through the series (the Horner form
// always stages it first).
template <int NC, int N>
constexpr bool exp_times_holds_b() {
  return !use_ps<N>() || N <= (NC == 2 ? 24 : 20);
}

// The form each instance takes: exp(X)*B into slot B (X, B as lane
// elements; the Horner form stages them itself) ...
//
// load_b() fills bv: in the Paterson-Stockmeyer form it runs once X is
// staged (B's loads then overlap the first product instead of adding to the
// start's memory wait), in the Horner form first.
template <int NC, int N, class LoadB = NoHook>
__device__ __forceinline__ void series_exp_times(const Lane<NC, N>& ln, int n, double2 (&X)[N], double2 bflag,
                                                 double2 (&bv)[N], LoadB&& load_b = NoHook{}) {
  if constexpr (use_ps<N>()) {
    ps_exp_times<NC, N>(ln, n, X, bflag, bv, load_b);
  } else {
    load_b();
    ln.stage(ln.slotA, X);
    ln.stage(ln.slotB, bv);
    horner_exp_times<NC, N>(ln, n, bflag, bv);
  }
}
// ... and exp(X) or log(1 + X) into slot A, X staged in slot B by the caller
// and also given as lane elements.
template <int NC, int N, bool IS_EXP>
__device__ __forceinline__ void series_fn(const Lane<NC, N>& ln, int n, const double2 (&X)[N]) {
  if constexpr (use_ps<N>())
    ps_series<NC, N, IS_EXP>(ln, n, X);
  else
    horner<NC, N, IS_EXP>(ln, n);
}

// reH of order k of the series in an LDS slot (QCDpt.cc:197-209):
// (M - M^dag)/2 minus Tr/NC, element (r, c) of this lane.
template <int NC, int N>
__device__ __forceinline__ double2 reh_elem(const Lane<NC, N>& ln, const double2* slot, int k) {
  constexpr int NE = NC * NC;
  const double2* m = slot + k * NE;
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

// This lane's element of every order of a series in an LDS slot.
template <int NC, int N>
__device__ __forceinline__ void unstage(const Lane<NC, N>& ln, const double2* slot, double2 (&X)[N]) {
  constexpr int NE = NC * NC;
#pragma unroll
  for (int k = 0; k < N; k++) X[k] = slot[k * NE + ln.e];
}

// Order k of a series lies k * NE * NSPT_TILE entries after its order 0.
// While a field stays within 2 GiB, one buffer descriptor, a 32-bit lane
// offset and a scalar byte offset per order reach every entry;
// NSPT_BIG_FIELDS builds address each lane's series with 64 bits (see the
// top of this file).
// Global loads/stores of a lane-owned series.  Orders >= nrt read as zero and
// are not stored.  Loads are unmasked (lanes of invalid links hold clamped,
// valid indices); stores are masked by ok.
template <int NC, int N>
__device__ __forceinline__ void gload(const SerField& f, long idx, int e, int nrt, bool ok, double2 (&X)[N]) {
  constexpr int NE = NC * NC;
  constexpr int plane = NE * NSPT_TILE;  // entries from order k to k+1 of a series
  (void)ok;
#if NSPT_BIG_FIELDS
  const double2* p = f.data + ser_index(idx, e, nrt, NE);
#pragma unroll
  for (int k = 0; k < N; k++) X[k] = (k < nrt) ? p[k * plane] : zero2();
#else
  const unsigned off = (unsigned)ser_index(idx, e, nrt, NE);
  const __amdgpu_buffer_rsrc_t r = rsrc(f.data);
#pragma unroll
  for (int k = 0; k < N; k++) X[k] = (k < nrt) ? ld_r(r, off, (unsigned)(k * plane * 16)) : zero2();
#endif
}
template <int NC, int N>
__device__ __forceinline__ void gstore(const SerField& f, long idx, int e, int nrt, bool ok, const double2 (&X)[N]) {
  constexpr int NE = NC * NC;
  constexpr int plane = NE * NSPT_TILE;
  if (!ok) return;
#if NSPT_BIG_FIELDS
  double2* p = f.data + ser_index(idx, e, nrt, NE);
#pragma unroll
  for (int k = 0; k < N; k++)
    if (k < nrt) p[k * plane] = X[k];
#else
  const unsigned off = (unsigned)ser_index(idx, e, nrt, NE);
#pragma unroll
  for (int k = 0; k < N; k++)
    if (k < nrt) st_u(f.data + k * plane, off, X[k]);
#endif
}
// f += X (own element), or f = X when first.
template <int NC, int N>
__device__ __forceinline__ void gaccum(const SerField& f, long idx, int e, int nrt, bool ok, const double2 (&X)[N],
                                       bool first) {
  constexpr int NE = NC * NC;
  constexpr int plane = NE * NSPT_TILE;
  if (!ok) return;
  double2* p = f.data + ser_index(idx, e, nrt, NE);
#pragma unroll
  for (int k = 0; k < N; k++)
    if (k < nrt) p[k * plane] = first ? X[k] : c_add(p[k * plane], X[k]);
}
__device__ __forceinline__ double2 gflag(const SerField& f, long idx, bool ok) {
  return (ok && f.flag) ? f.flag[idx] : zero2();
}
// The flag of series idx in every lane, for callers whose idx is a valid
// series index in every lane (lanes past the block's links read a clamped
// site and never use their flag): an unconditional load.  gflag's per-lane
// condition makes a divergent join that waits for the load at once, and,
// vector loads completing in order, for every load issued before it -- the
// next operands a kernel prefetches during a product included.
__device__ __forceinline__ double2 gflag_all(const SerField& f, long idx) {
  return f.flag ? f.flag[idx] : zero2();
}

}  // inline namespace NSPT_VARIANT
}  // namespace nspt
