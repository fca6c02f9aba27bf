// nspt_kernels.hpp -- per-link / per-site kernels of the NSPT step, templated
// on the gauge group (NC = 3 or 2) and the compiled order count N (>= the
// runtime order; orders above the runtime order are staged as zero and never
// stored, which leaves the lower orders unchanged because a truncated product
// only feeds orders upward).
//
// Link index on the device: idx = mu*V + site (site lexicographic with x3
// fastest, lattice.h:37).  One wave = one workgroup = LPB links of the same mu.
// Every product stages both operands in LDS first (loads of A and B from HBM
// issued together), then runs from LDS (see nspt_device.hpp).
#pragma once
#include "nspt_device.hpp"
#include "nspt_internal.hpp"
#include "nspt_table.hpp"

#ifndef NSPT_DRIFT_P_PREFETCH_U
#define NSPT_DRIFT_P_PREFETCH_U 1
#endif
#ifndef NSPT_QT_DRIFT_P
#define NSPT_QT_DRIFT_P NSPT_QT_DRIFT
#endif

namespace nspt {
inline namespace NSPT_VARIANT {

__device__ __forceinline__ int site_shift(const Geo& g, int s, int d, int delta) {
  const int x = (s / g.st[d]) % g.L[d];
  int xn = x + delta;
  xn += (xn < 0) ? g.L[d] : 0;
  xn -= (xn >= g.L[d]) ? g.L[d] : 0;
  return s + (xn - x) * g.st[d];
}

// Shift by delta (either sign, |delta| <= L_d) along d.
__device__ __forceinline__ int site_move(const Geo& g, int s, int d, int delta) {
  const int x = (s / g.st[d]) % g.L[d];
  int xn = x + delta;
  xn += (xn < 0) ? g.L[d] : 0;
  xn -= (xn >= g.L[d]) ? g.L[d] : 0;
  return s + (xn - x) * g.st[d];
}

// Shift by delta >= 0 sites along d, any number of times around the lattice.
__device__ __forceinline__ int site_shift_far(const Geo& g, int s, int d, int delta) {
  const int x = (s / g.st[d]) % g.L[d];
  return s + ((x + delta) % g.L[d] - x) * g.st[d];
}

// ------------------------------------------------------------ twisted boundaries
// A neighbour reached from a site carries a crossing code: 3 bits per
// direction d holding n_d + 2, n_d the net number of times the step crossed
// the boundary of d (|n_d| <= 2).  TW0 = no crossing.  On a twisted lattice
// the link read there is the twisted copy W U W^dag, W = prod_d W_d^(n_d)
// (the halo copies of twist_boundary, nspt.cc:163-256: forward W A W^dag,
// backward W^dag A W).
constexpr int TW0 = 2 | (2 << 3) | (2 << 6) | (2 << 9);

// site_move / site_shift_far that also count the boundary crossings.
__device__ __forceinline__ int site_move_tw(const Geo& g, int s, int d, int delta, int& code) {
  const int x = (s / g.st[d]) % g.L[d];
  int xn = x + delta, n = 0;
  if (xn < 0) {
    xn += g.L[d];
    n = -1;
  } else if (xn >= g.L[d]) {
    xn -= g.L[d];
    n = 1;
  }
  code += n * (1 << (3 * d));
  return s + (xn - x) * g.st[d];
}
__device__ __forceinline__ int site_shift_far_tw(const Geo& g, int s, int d, int delta, int& code) {
  const int x = (s / g.st[d]) % g.L[d];
  code += ((x + delta) / g.L[d]) << (3 * d);
  return s + ((x + delta) % g.L[d] - x) * g.st[d];
}

// One factor of a row walk: row i of M times W_d (inv: W_d^dag) is row
// i' of the product with phase m.
// The table is wave-uniform: its words are read with scalar loads and the
// lane's entry picked by selects.
template <int NC>
__device__ __forceinline__ void tw_step(const TwistTab* tw, int d, bool inv, int& i, double2& m) {
  d = uniform(d);
  const unsigned perm = tw->perm[d];
  const double2 p0 = tw->ph[d][0], p1 = tw->ph[d][1], p2 = NC > 2 ? tw->ph[d][2] : p1;
  const int k = inv ? (int)((perm >> (8 + 2 * i)) & 3) : i;
  const double2 w = k == 0 ? p0 : (k == 1 ? p1 : p2);
  m = c_mul(m, inv ? c_conj(w) : w);
  i = inv ? k : (int)((perm >> (2 * i)) & 3);
}
// Element (r, c) of the twisted copy W U W^dag is ph times element
// (pe / NC, pe % NC) of U: W monomial, W[r][pi(r)] = m_r, so
// (W U W^dag)[r][c] = m_r conj(m_c) U[pi(r)][pi(c)].
template <int NC>
__device__ __forceinline__ void twist_elem(const Geo& g, int code, int r, int c, int& pe, double2& ph) {
  double2 mr = one2(), mc = one2();
  if (code != TW0) {
    for (int d = 0; d < 4; d++) {
      const int n = ((code >> (3 * d)) & 7) - 2;
      for (int t = 0; t < n || t < -n; t++) {
        tw_step<NC>(g.tw, d, n < 0, r, mr);
        tw_step<NC>(g.tw, d, n < 0, c, mc);
      }
    }
  }
  pe = r * NC + c;
  ph = c_mul(mr, c_conj(mc));
}

// Block order.  The dispatcher hands consecutive workgroups to the 8 XCDs in
// turn; block_order() renumbers them so that each XCD runs one contiguous
// range of logical blocks (neighbour links then meet in that XCD's L2), and
// link kernels take the direction as the fastest logical index so that the
// four links of a site block, which share most staple links, run together.
#ifndef NSPT_BLOCK_ORDER
#define NSPT_BLOCK_ORDER 2
#endif
// Block p of n (a replica-table launch passes the replica's own count; its
// grid's x extent is a multiple of 8, so x = p & 7 is still the XCD).
__device__ __forceinline__ int block_order(int p, int n) {
#if NSPT_BLOCK_ORDER >= 2
  const int q = n >> 3, r = n & 7, x = p & 7;
  return x * q + min(x, r) + (p >> 3);
#else
  (void)n;
  return p;
#endif
}
__device__ __forceinline__ int block_order() { return block_order(blockIdx.x, gridDim.x); }

// Link block of this workgroup: direction mu, first index site0 in the list
// of its sites (the lattice, or a checkerboard subset of count = V/2 sites),
// and the index of its per-block partial sums (mu-major, as the reductions
// expect: a subset pass (mu, parity) owns rows (2 mu + parity) nb_half + b).
struct LinkBlock {
  int mu, site0, part, count;
};
template <int LPB>
__device__ __forceinline__ LinkBlock link_block_at(const Geo& g, int b) {
  if (g.sub) {
    const int mu = (g.sub - 1) & 3, par = (g.sub - 1) >> 2;
    return {mu, b * LPB, (2 * mu + par) * g.nb_half + b, g.V / 2};
  }
#if NSPT_BLOCK_ORDER >= 1
  const int mu = b & 3, sb = b >> 2;
#else
  const int mu = b / g.nb_mu, sb = b % g.nb_mu;
#endif
  return {mu, sb * LPB, mu * g.nb_mu + sb, g.V};
}
template <int LPB>
__device__ __forceinline__ LinkBlock link_block(const Geo& g) {
  return link_block_at<LPB>(g, block_order());
}

// Site of entry k of a block's site list: k itself, or the k-th site of the
// subset's parity (with L_3 even, sites 2k and 2k+1 share x0, x1, x2 and
// have opposite parities).
__device__ __forceinline__ int link_site(const Geo& g, int k) {
  if (!g.sub) return k;
  const int par = (g.sub - 1) >> 2;
  const int s2 = 2 * k;
  int r = s2 / g.L[3];
  const int x2 = r % g.L[2];
  r /= g.L[2];
  const int x1 = r % g.L[1];
  const int x0 = r / g.L[1];
  return s2 + (((x0 + x1 + x2) ^ par) & 1);
}

// Lane-owned element of series idx of field f, all orders, into an LDS slot.
// Loads are unmasked: lanes of invalid links carry a clamped, valid index.
// With dag, the dagger of the series is staged (element (r, c) loaded,
// conj written at (c, r)).
// The loads go out in chunks of STAGE_CHUNK orders with no branch between
// them (an order at or above nrt reads plane 0 and is written as zero), so
// a chunk waits for one memory latency; a per-order "load if k < nrt, then
// write" compiles to a branch per order and waits for every load in turn.
// (CH = 1 keeps the per-order form, which measured faster where the
// staging follows a long product with every register in use: the gauge
// transform above 24 orders.)
#ifndef NSPT_STAGE_CHUNK
#define NSPT_STAGE_CHUNK 16
#endif
template <int NC, int N, int CH = NSPT_STAGE_CHUNK>
__device__ __forceinline__ void stage_global(const Lane<NC, N>& ln, double2* slot, const SerField& f, long idx,
                                             int nrt, bool dag = false) {
  constexpr int NE = NC * NC;
  constexpr int plane = NE * NSPT_TILE;
  if (!ln.lds_ok) return;
  const int pos = dag ? (ln.c * NC + ln.r) : ln.e;
  const double sy = dag ? -1.0 : 1.0;
#if NSPT_BIG_FIELDS
  const double2* p = f.data + ser_index(idx, ln.e, nrt, NE);
#else
  const unsigned off = (unsigned)ser_index(idx, ln.e, nrt, NE);
  const __amdgpu_buffer_rsrc_t r = rsrc(f.data);
#endif
#pragma unroll
  for (int k0 = 0; k0 < N; k0 += CH) {
    double2 v[CH];
#pragma unroll
    for (int i = 0; i < CH; i++) {
      const int k = k0 + i;
      if (k < N) {
        const int kk = uniform(k < nrt ? k : 0);
#if NSPT_BIG_FIELDS
        v[i] = p[kk * plane];
#else
        v[i] = ld_r(r, off, (unsigned)(kk * plane * 16));
#endif
      }
    }
#pragma unroll
    for (int i = 0; i < CH; i++) {
      const int k = k0 + i;
      if (k < N) {
        const double2 w = (k < nrt) ? v[i] : zero2();
        slot[k * NE + pos] = make_double2(w.x, sy * w.y);
      }
    }
  }
}

// Loads and stores of a link reached with crossing code `code`: on a
// twisted lattice the lane reads (writes) the element of the stored link
// that its element of the twisted copy comes from (goes to), times the
// phase (its conjugate).
template <int NC, int N>
__device__ __forceinline__ void gload_tw(const Geo& g, const Lane<NC, N>& ln, const SerField& f, long idx, int code,
                                         int nrt, bool ok, double2 (&X)[N]) {
  // (only waves with a lane across a twisted boundary take the twisted path)
  const bool tw = g.twist && __any(code != TW0);
  int pe = ln.e;
  double2 ph = one2();
  if (tw) twist_elem<NC>(g, code, ln.r, ln.c, pe, ph);
  gload<NC, N>(f, idx, pe, nrt, ok, X);
  if (tw) {
#pragma unroll
    for (int k = 0; k < N; k++) X[k] = c_mul(ph, X[k]);
  }
}
template <int NC, int N>
__device__ __forceinline__ void gstore_tw(const Geo& g, const Lane<NC, N>& ln, const SerField& f, long idx, int code,
                                          int nrt, bool ok, double2 (&X)[N]) {
  int pe = ln.e;
  double2 ph = one2();
  if (g.twist && __any(code != TW0)) {
    twist_elem<NC>(g, code, ln.r, ln.c, pe, ph);
#pragma unroll
    for (int k = 0; k < N; k++) X[k] = c_mul(c_conj(ph), X[k]);
  }
  gstore<NC, N>(f, idx, pe, nrt, ok, X);
}
// stage_global of a link reached with crossing code `code`.
template <int NC, int N>
__device__ __forceinline__ void stage_global_tw(const Geo& g, const Lane<NC, N>& ln, double2* slot, const SerField& f,
                                                long idx, int code, int nrt, bool dag = false) {
  if (!g.twist || !__any(code != TW0)) {
    stage_global(ln, slot, f, idx, nrt, dag);
    return;
  }
  double2 X[N];
  gload_tw(g, ln, f, idx, code, nrt, true, X);
  ln.stage(slot, X, dag);
}

// Sum over the valid link slots of the block of element (k, e) of slot A.
// Deterministic (ascending slots).
template <int NC, int N>
__device__ __forceinline__ double2 slot_sum(const double2* slotA0, int nvalid, int k, int e) {
  constexpr int NE = NC * NC;
  double2 s = zero2();
  for (int jj = 0; jj < nvalid; jj++) s = c_add(s, slotA0[jj * Lane<NC, N>::SLOT + k * NE + e]);
  return s;
}

// End of a drift evaluation, with U_mu(x) in slot A and W in slot B (flag
// wf; U's flag uflag): the observables of the block (op, may be null: plaquette
// sums Tr W per order, norm and Tr U of U_mu(x)), the improved-action term
// W = c0 W + c1 U R, and D = tau_g reH(W) (mode 1: the RK2 combination with
// D1).  mode 2 stops after the observables.
template <int NC, int N>
__device__ __forceinline__ void drift_finish(const Lane<NC, N>& ln, double2* lds, double2* flagbuf, int nvalid,
                                             int nrt, int M, double* op, double2 wf, double2 uflag, const SerField& R,
                                             double c1, long idx, bool ok, double tau_g, int mode, const SerField& D,
                                             const SerField& D1, double rk2c) {
  constexpr int NE = NC * NC;
  using LN = Lane<NC, N>;
  if (op) {
    const int lane = threadIdx.x;
    // norm_k = sum Re Tr(U_k U_k^dag), trU_k = sum Tr U_k over U_mu(x)
    if (lane < nrt) {
      double nr = 0.0;
      double2 tr = zero2();
      for (int jj = 0; jj < nvalid; jj++) {
        const double2* m = lds + jj * LN::SLOT + lane * NE;
#pragma unroll
        for (int q = 0; q < NE; q++) nr += m[q].x * m[q].x + m[q].y * m[q].y;
#pragma unroll
        for (int q = 0; q < NC; q++) tr = c_add(tr, m[q * NC + q]);
      }
      op[2 * (nrt + 1) + lane] = nr;
      op[3 * nrt + 2 + 2 * lane] = tr.x;
      op[3 * nrt + 2 + 2 * lane + 1] = tr.y;
    }
    if (ln.lds_ok && ln.e == 0) flagbuf[ln.j] = wf;
    __syncthreads();
    // plaquette sums Tr W_k (k = 0 from the flags)
    if (lane <= nrt) {
      double2 s = zero2();
      for (int jj = 0; jj < nvalid; jj++) {
        if (lane == 0) {
          s = c_add(s, c_scale((double)NC, flagbuf[jj]));
        } else {
          const double2* m = lds + LN::LPB * LN::SLOT + jj * LN::SLOT + (lane - 1) * NE;
#pragma unroll
          for (int q = 0; q < NC; q++) s = c_add(s, m[q * NC + q]);
        }
      }
      op[2 * lane] = s.x;
      op[2 * lane + 1] = s.y;
    }
  }
  if (mode == 2) return;

  if (R.data) {
    // improved action (nspt.cc:405-412): W = c0 U Sigma + c1 U R, R the sum
    // of the rectangle staples, c0 = 1 - 8 c1; U_mu(x) is still in slot A
    double2 W1[N], T[N];
    unstage(ln, ln.slotB, W1);
    __syncthreads();
    stage_global(ln, ln.slotB, R, idx, nrt);
    const double2 rf = gflag(R, idx, ok);
#pragma unroll
    for (int k = 0; k < N; k++) T[k] = zero2();
    __syncthreads();
    mul<NC, N>(ln, uflag, rf, true, T, M, false, nullptr, false, 1.0);
    __syncthreads();
    const double c0 = 1.0 - 8.0 * c1;
#pragma unroll
    for (int k = 0; k < N; k++) W1[k] = c_add(c_scale(c0, W1[k]), c_scale(c1, T[k]));
    ln.stage(ln.slotB, W1);
    __syncthreads();
  }

  // D = tau_g reH(W)
  double2 Out[N];
#pragma unroll
  for (int k = 0; k < N; k++) Out[k] = c_scale(tau_g, reh_elem(ln, ln.slotB, k));
  if (mode == 1) {
    double2 D1v[N];
    gload<NC, N>(D1, idx, ln.e, nrt, ok, D1v);
#pragma unroll
    for (int k = 0; k < N; k++) Out[k] = c_scale(0.5, c_add(D1v[k], Out[k]));
#pragma unroll
    for (int k = N - 1; k >= 2; k--) Out[k] = c_add(Out[k], c_scale(rk2c, Out[k - 2]));
  }
  gstore<NC, N>(D, idx, ln.e, nrt, ok, Out);
}

// ------------------------------------------------------------ drift
// D = tau_g * reH(U_mu(x) * sum_{nu != mu} staple_{mu nu}(x))
// (nspt.cc:383-417, staple QCDenvNODEpt_test.h:554-557).
// mode 0: store D.  mode 1: RK2 corrector, store (D1 + D)/2 * (1 + rk2c x^2).
// mode 2: observables only.  obs_part (may be null): per block
// [plaq (N+1) cplx][norm N][trU N cplx] of the block's mu, unnormalised.
template <int NC, int N>
__device__ __forceinline__ void k_drift_body(int blk_, int nblk_, const Geo& g, SerField U, SerField D, SerField D1,
                                                                 double tau_g, double rk2c, int mode,
                                                                 double* obs_part, SerField R, double c1) {
  constexpr int NE = NC * NC;
  using LN = Lane<NC, N>;
  __shared__ double2 lds[LN::LDS];
  __shared__ double2 flagbuf[LN::LPB];
  LN ln(lds);
  const int nrt = g.nrt;
  const int M = min(N, nrt) - 1;
  const LinkBlock lb = link_block_at<LN::LPB>(g, block_order(blk_, nblk_));
  const int mu = lb.mu;
  const int site0 = lb.site0;
  const int nvalid = min(LN::LPB, lb.count - site0);
  const int site = link_site(g, min(site0 + ln.j, lb.count - 1));
  const bool ok = ln.lds_ok && (site0 + ln.j) < lb.count;
  const long V = g.V;
  const long idx = mu * V + site;

  double2 Sig[N];
#pragma unroll
  for (int k = 0; k < N; k++) Sig[k] = zero2();
  double2 sigf = zero2(), t1f = zero2(), t2f = zero2(), wf = zero2(), uflag = zero2();

  // 12 staple products (4 per plane nu != mu) and W = U_mu(x) Sigma, one
  // product per iteration so that the product body is emitted once:
  //   kind 0: T1   = U_nu(x)      U_mu(x+nu)   -> T1^dag into slot B
  //   kind 1: up   = U_nu(x+mu)   T1^dag       -> Sigma +=
  //   kind 2: T2   = U_mu(x-nu)   U_nu(x-nu+mu) -> T2^dag into slot A
  //   kind 3: down = T2^dag       U_nu(x-nu)   -> Sigma +=
  //   i = 12: W    = U_mu(x)      Sigma        -> slot B
  // T1^dag and T2^dag are formed in place in the slot that the next product
  // reads.  The links of product i+1 are loaded into PA/PB while product i
  // runs from LDS.
  int cpmu = TW0;
  const int xpmu = site_move_tw(g, site, mu, 1, cpmu);
  auto kind_of = [](int i) { return (i == 12) ? 4 : (i & 3); };
  // links of product i and their boundary-crossing codes
  auto links_of = [&](int i, long& ia, long& ib, int& ca, int& cb) {
    const int p = i >> 2, kind = kind_of(i);
    const int nu = (p < mu) ? p : p + 1;
    int cpnu = TW0, cmnu = TW0;
    const int xpnu = site_move_tw(g, site, nu, 1, cpnu);
    const int xmnu = site_move_tw(g, site, nu, -1, cmnu);
    ia = idx;
    ib = idx;
    ca = TW0;
    cb = TW0;
    if (kind == 0) {
      ia = nu * V + site;
      ib = mu * V + xpnu;
      cb = cpnu;
    } else if (kind == 1) {
      ia = nu * V + xpmu;
      ca = cpmu;
    } else if (kind == 2) {
      ia = mu * V + xmnu;
      ca = cmnu;
      cb = cmnu;
      ib = nu * V + site_move_tw(g, xmnu, mu, 1, cb);
    } else if (kind == 3) {
      ib = nu * V + xmnu;
      cb = cmnu;
    }
  };
  // Up to N = 24 the next product's links are prefetched into registers;
  // above that the registers are needed for Sigma and the product, and the
  // links are staged straight from HBM.
  constexpr bool PF = N <= 24;
  constexpr int NP = PF ? N : 1;
  double2 PA[NP], PB[NP], paf = zero2(), pbf = zero2();
  auto fetch = [&](int i) {
    long ia, ib;
    int ca, cb;
    links_of(i, ia, ib, ca, cb);
    const int kind = kind_of(i);
    if (kind != 3) {
      if constexpr (PF) gload_tw<NC, N>(g, ln, U, ia, ca, nrt, ok, PA);
      paf = gflag_all(U, ia);
    }
    if (kind == 0 || kind == 2 || kind == 3) {
      if constexpr (PF) gload_tw<NC, N>(g, ln, U, ib, cb, nrt, ok, PB);
      pbf = gflag_all(U, ib);
    }
  };
  fetch(0);
  for (int i = 0; i < 13; i++) {
    asm volatile("" ::: "memory");
    const int kind = kind_of(i);
    if constexpr (PF) {
      if (kind != 3) ln.stage(ln.slotA, PA);
      if (kind == 0 || kind == 2 || kind == 3) ln.stage(ln.slotB, PB);
    } else {
      long ia, ib;
      int ca, cb;
      links_of(i, ia, ib, ca, cb);
      if (kind != 3) stage_global_tw(g, ln, ln.slotA, U, ia, ca, nrt);
      if (kind == 0 || kind == 2 || kind == 3) stage_global_tw(g, ln, ln.slotB, U, ib, cb, nrt);
    }
    if (kind == 4) ln.stage(ln.slotB, Sig);
    const double2 a0 = (kind == 3) ? c_conj(t2f) : paf;
    if (kind == 4) uflag = a0;
    const double2 b0 = (kind == 1) ? c_conj(t1f) : (kind == 4) ? sigf : pbf;
    __syncthreads();
    if (i + 1 < 13) fetch(i + 1);
    const bool to_lds = (kind == 0 || kind == 2 || kind == 4);
    double2* dst = (kind == 2) ? ln.slotA : ln.slotB;
    const double2 f = mul<NC, N, 1, false, NSPT_QT_DRIFT>(ln, a0, b0, true, Sig, M, to_lds, dst, kind != 4, 1.0);
    if (kind == 0)
      t1f = f;
    else if (kind == 2)
      t2f = f;
    else if (kind == 4)
      wf = f;
    else
      sigf = c_add(sigf, f);
    __syncthreads();
  }
  // slot A: U_mu(x); slot B: W
  drift_finish<NC, N>(ln, lds, flagbuf, nvalid, nrt, M,
                      obs_part ? obs_part + (long)lb.part * (2 * (nrt + 1) + 3 * nrt) : nullptr, wf, uflag, R, c1,
                      idx, ok, tau_g, mode, D, D1, rk2c);
}

template <int NC, int N>
__global__ __launch_bounds__(64, NSPT_WAVES_PER_EU) void k_drift(Geo g, SerField U, SerField D, SerField D1,
                                                                 double tau_g, double rk2c, int mode,
                                                                 double* obs_part, SerField R, double c1) {
  k_drift_body<NC, N>(blockIdx.x, gridDim.x, g, U, D, D1, tau_g, rk2c, mode, obs_part, R, c1);
}
template <int NC, int N>
__global__ __launch_bounds__(64, NSPT_WAVES_PER_EU) void k_drift_tab(const DriftArgs* __restrict__ t) {
  const DriftArgs& a = t[blockIdx.y];
  if ((int)blockIdx.x >= a.nblk) return;
  k_drift_body<NC, N>(blockIdx.x, a.nblk, a.g, a.U, a.D, a.D1, a.tau_g, a.rk2c, a.mode, a.obs_part, a.R, a.c1);
}

// ------------------------------------------------------------ drift from the plaquette field
// The same drift without forming the staple sum: W = U_mu(x) Sigma_mu(x) is
// a sum over nu != mu of the two plaquettes through the link, rotated to
// start there,
//   W = sum_nu P_mu,nu(x) + U_nu(x-nu)^dag P_mu,nu(x-nu)^dag U_nu(x-nu),
//   P_mu,nu(x) = U_mu(x) U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag,
// and P_nu,mu = P_mu,nu^dag, so the six plaquettes of a site (3 products
// each) serve all four of its links: 4.5 + 6 products per link instead of
// the 13 of k_drift.  The rotation of the lower plaquette equals the
// reference's staple term U_mu(x) U_nu(x+mu-nu)^dag U_mu(x-nu)^dag U_nu(x-nu)
// because U_nu(x-nu)^dag U_nu(x-nu) = 1, which the field satisfies order by
// order: the cold start is 1, and every update and gauge transform
// multiplies it by exponentials of antihermitian series (to rounding).
//
// k_plaq_field: P of plane p = (mu < nu) at every site, stored at series
// index p V + site, planes (0,1) (0,2) (0,3) (1,2) (1,3) (2,3).
__device__ __forceinline__ int plane_index(int a, int b) {
  // a < b
  return a == 0 ? b - 1 : (a == 1 ? b + 1 : 5);
}
template <int NC, int N>
__device__ __forceinline__ void k_plaq_field_body(int blk_, int nblk_, const Geo& g, SerField U, SerField P) {
  using LN = Lane<NC, N>;
  __shared__ double2 lds[LN::LDS];
  LN ln(lds);
  const int nrt = g.nrt;
  const int M = min(N, nrt) - 1;
  const int b = block_order(blk_, nblk_);
  const int plane = b % 6, sb = b / 6;
  const int mu = plane < 3 ? 0 : (plane < 5 ? 1 : 2);
  const int nu = plane < 3 ? plane + 1 : (plane < 5 ? plane - 1 : 3);
  const int site0 = sb * LN::LPB;
  const int site = min(site0 + ln.j, g.V - 1);
  const bool ok = ln.lds_ok && (site0 + ln.j) < g.V;
  const long V = g.V;
  int cpn = TW0, cpm = TW0;
  const int xpnu = site_move_tw(g, site, nu, 1, cpn);
  const int xpmu = site_move_tw(g, site, mu, 1, cpm);
  // product 0: C = U_nu(x) U_mu(x+nu) -> C^dag in place in slot B
  // product 1: T = U_nu(x+mu) C^dag   -> in place in slot B
  // product 2: P = U_mu(x) T          -> registers (PB, free after staging)
  // one product per iteration (the product body is emitted once); the next
  // left factor is loaded into PA while the current product runs
  double2 PA[N], PB[N];
  gload<NC, N>(U, nu * V + site, ln.e, nrt, ok, PA);
  gload_tw<NC, N>(g, ln, U, mu * V + xpnu, cpn, nrt, ok, PB);
  double2 fa = gflag_all(U, nu * V + site), fb = gflag_all(U, mu * V + xpnu);
  ln.stage(ln.slotB, PB);
#pragma unroll
  for (int k = 0; k < N; k++) PB[k] = zero2();
  for (int i = 0; i < 3; i++) {
    asm volatile("" ::: "memory");
    ln.stage(ln.slotA, PA);
    const double2 a0 = fa;
    __syncthreads();
    if (i < 2) {
      const long il = (i == 0) ? nu * V + xpmu : mu * V + site;
      gload_tw<NC, N>(g, ln, U, il, i == 0 ? cpm : TW0, nrt, ok, PA);
      fa = gflag_all(U, il);
    }
    const double2 f = mul<NC, N, 1, false, NSPT_QT_DRIFT>(ln, a0, fb, true, PB, M, i < 2, ln.slotB, i == 0, 1.0);
    fb = (i == 0) ? c_conj(f) : f;
    __syncthreads();
  }
  const double2 pf = fb;
  const long ip = (long)plane * V + site;
  gstore<NC, N>(P, ip, ln.e, nrt, ok, PB);
  if (ok && ln.e == 0) P.flag[ip] = pf;
}

template <int NC, int N>
__global__ __launch_bounds__(64, NSPT_WAVES_PER_EU) void k_plaq_field(Geo g, SerField U, SerField P) {
  k_plaq_field_body<NC, N>(blockIdx.x, gridDim.x, g, U, P);
}
template <int NC, int N>
__global__ __launch_bounds__(64, NSPT_WAVES_PER_EU) void k_plaq_field_tab(const PlaqArgs* __restrict__ t) {
  const PlaqArgs& a = t[blockIdx.y];
  if ((int)blockIdx.x >= a.nblk) return;
  k_plaq_field_body<NC, N>(blockIdx.x, a.nblk, a.g, a.U, a.P);
}

// k_drift_p: the drift of k_drift (same modes, observables and improved
// action) from the plaquette field of U.  Per nu: t = U_nu(x-nu)^dag X
// (X = P_mu,nu(x-nu)^dag) in place in slot A, then Sigma += t U_nu(x-nu);
// the plaquettes at x are added as loaded.
template <int NC, int N>
__device__ __forceinline__ void k_drift_p_body(int blk_, int nblk_, const Geo& g, SerField U, SerField P, SerField D,
                                                                   SerField D1, double tau_g, double rk2c, int mode,
                                                                   double* obs_part, SerField R, double c1) {
  constexpr int NE = NC * NC;
  using LN = Lane<NC, N>;
  __shared__ double2 lds[LN::LDS];
  __shared__ double2 flagbuf[LN::LPB];
  LN ln(lds);
  const int nrt = g.nrt;
  const int M = min(N, nrt) - 1;
  const LinkBlock lb = link_block_at<LN::LPB>(g, block_order(blk_, nblk_));
  const int mu = lb.mu;
  const int site0 = lb.site0;
  const int nvalid = min(LN::LPB, lb.count - site0);
  const int site = min(site0 + ln.j, lb.count - 1);
  const bool ok = ln.lds_ok && (site0 + ln.j) < lb.count;
  const long V = g.V;
  const long idx = mu * V + site;
  const int etr = ln.c * NC + ln.r;  // element of the transpose

  // the plaquettes at x: P_mu,nu(x), or P_nu,mu(x)^dag for nu < mu (the
  // first loaded straight into Sigma, the other two after the products,
  // into the prefetch registers; staging them through LDS before the
  // products measured 5% slower at 16^4 N=20, and loading the RK2 D1 at the
  // start of the finish 2.5% slower: the register budget decides)
  auto at_x = [&](int p, bool& dg) {
    const int nu = (p < mu) ? p : p + 1;
    dg = nu < mu;
    return (long)plane_index(min(mu, nu), max(mu, nu)) * V + site;
  };
  double2 Sig[N];
  double2 sigf;
  {
    bool dg;
    const long ip = at_x(0, dg);
    gload<NC, N>(P, ip, dg ? etr : ln.e, nrt, ok, Sig);
    sigf = gflag(P, ip, ok);
    if (dg) {
      sigf = c_conj(sigf);
#pragma unroll
      for (int k = 0; k < N; k++) Sig[k] = c_conj(Sig[k]);
    }
  }
  // the plaquettes at x - nu through the link.  Up to N = 20 (24 for
  // SU(2), whose smaller products leave more registers) the next plaquette
  // and link are prefetched into registers; above that the registers are
  // needed for Sigma and the product, and they are staged straight from HBM.
  constexpr bool PF = N <= (NC == 2 ? 24 : 20);
  constexpr int NP = PF ? N : 1;
  // with PFU the link U_nu(x-nu) is prefetched too; without it, it is
  // staged straight from HBM (80 fewer live VGPRs during the products)
  constexpr bool PFU = PF && NSPT_DRIFT_P_PREFETCH_U;
  constexpr int NPU = PFU ? N : 1;
  double2 PX[NP], PU[NPU];
  int cm = TW0;
  long ix, iu;
  bool xdag;
  auto locate = [&](int p) {
    const int nu = (p < mu) ? p : p + 1;
    cm = TW0;
    const int xm = site_move_tw(g, site, nu, -1, cm);
    // P_mu,nu(x-nu)^dag: the stored P_mu,nu daggered (nu > mu), or the
    // stored P_nu,mu itself (nu < mu)
    xdag = nu > mu;
    ix = (long)plane_index(min(mu, nu), max(mu, nu)) * V + xm;
    iu = (long)nu * V + xm;
  };
  locate(0);
  if constexpr (PF) gload_tw<NC, N>(g, ln, P, ix, cm, nrt, ok, PX);
  if constexpr (PFU) gload_tw<NC, N>(g, ln, U, iu, cm, nrt, ok, PU);
  double2 xf = gflag_all(P, ix), uf = gflag_all(U, iu);
  bool xd = xdag;
  double2 tf = zero2(), ufc = zero2();
  // one product per iteration (the product body is emitted once):
  //   even i: slot A <- X^dag = P_mu,nu(x-nu) (the stored P, or the dagger
  //           of the stored P_nu,mu), slot B <- U_nu(x-nu); X^dag U is
  //           retired daggered in place into slot A: t = U^dag X
  //   odd i:  Sigma += t U_nu(x-nu)
  // the next plaquette and link (last: U_mu(x)) load during the even products
  for (int i = 0; i < 6; i++) {
    asm volatile("" ::: "memory");
    const bool even = (i & 1) == 0;
    double2 a0 = tf;
    if (even) {
      if constexpr (PF)
        ln.stage(ln.slotA, PX, !xd);
      else
        stage_global_tw(g, ln, ln.slotA, P, ix, cm, nrt, !xd);
      if constexpr (PFU)
        ln.stage(ln.slotB, PU);
      else
        stage_global_tw(g, ln, ln.slotB, U, iu, cm, nrt);
      ufc = uf;
      a0 = xd ? xf : c_conj(xf);
    }
    __syncthreads();
    if (even) {
      if (i + 2 < 6) {
        locate(i / 2 + 1);
        if constexpr (PF) gload_tw<NC, N>(g, ln, P, ix, cm, nrt, ok, PX);
        if constexpr (PFU) gload_tw<NC, N>(g, ln, U, iu, cm, nrt, ok, PU);
        xf = gflag_all(P, ix);
        uf = gflag_all(U, iu);
        xd = xdag;
      } else {
        if constexpr (PFU) gload<NC, N>(U, idx, ln.e, nrt, ok, PU);  // U_mu(x), for the observables
        uf = gflag_all(U, idx);
      }
    }
    const double2 f = mul<NC, N, 1, false, NSPT_QT_DRIFT_P>(ln, a0, ufc, true, Sig, M, even, ln.slotA, even, 1.0);
    if (even)
      tf = c_conj(f);
    else
      sigf = c_add(sigf, f);
    __syncthreads();
  }
  // U_mu(x) into slot A; the other two plaquettes at x into Sigma
  if constexpr (PFU) {
    ln.stage(ln.slotA, PU);
    bool d1, d2;
    const long i1 = at_x(1, d1), i2 = at_x(2, d2);
    gload<NC, N>(P, i1, d1 ? etr : ln.e, nrt, ok, PX);
    gload<NC, N>(P, i2, d2 ? etr : ln.e, nrt, ok, PU);
    const double2 f1 = gflag(P, i1, ok), f2 = gflag(P, i2, ok);
    sigf = c_add(sigf, c_add(d1 ? c_conj(f1) : f1, d2 ? c_conj(f2) : f2));
#pragma unroll
    for (int k = 0; k < N; k++)
      Sig[k] = c_add(Sig[k], c_add(d1 ? c_conj(PX[k]) : PX[k], d2 ? c_conj(PU[k]) : PU[k]));
  } else {
    for (int q = 1; q < 3; q++) {
      bool dq;
      const long iq = at_x(q, dq);
      stage_global(ln, ln.slotB, P, iq, nrt, dq);
      const double2 fq = gflag(P, iq, ok);
      sigf = c_add(sigf, dq ? c_conj(fq) : fq);
      __syncthreads();
#pragma unroll
      for (int k = 0; k < N; k++) Sig[k] = c_add(Sig[k], ln.slotB[k * NE + ln.e]);
      __syncthreads();
    }
    stage_global(ln, ln.slotA, U, idx, nrt);
  }
  ln.stage(ln.slotB, Sig);
  __syncthreads();
  // slot A: U_mu(x); slot B: W
  drift_finish<NC, N>(ln, lds, flagbuf, nvalid, nrt, M,
                      obs_part ? obs_part + (long)lb.part * (2 * (nrt + 1) + 3 * nrt) : nullptr, sigf, uf, R, c1,
                      idx, ok, tau_g, mode, D, D1, rk2c);
}

template <int NC, int N>
__global__ __launch_bounds__(64, NSPT_WAVES_PER_EU) void k_drift_p(Geo g, SerField U, SerField P, SerField D,
                                                                   SerField D1, double tau_g, double rk2c, int mode,
                                                                   double* obs_part, SerField R, double c1) {
  k_drift_p_body<NC, N>(blockIdx.x, gridDim.x, g, U, P, D, D1, tau_g, rk2c, mode, obs_part, R, c1);
}
template <int NC, int N>
__global__ __launch_bounds__(64, NSPT_WAVES_PER_EU) void k_drift_p_tab(const DriftPArgs* __restrict__ t) {
  const DriftPArgs& a = t[blockIdx.y];
  if ((int)blockIdx.x >= a.nblk) return;
  k_drift_p_body<NC, N>(blockIdx.x, a.nblk, a.g, a.U, a.P, a.D, a.D1, a.tau_g, a.rk2c, a.mode, a.obs_part, a.R, a.c1);
}

// ------------------------------------------------------------ update
// Uout = exp(D - stau*eta at order 1) * Ubase  (nspt.cc:420-430, exp QCDpt.h:200-220);
// optionally L = reH(log(Uout)) (QCDpt.h:176-198, 197-209) and per-block sums
// of L for the zero mode (nspt.cc:432-439).  Uout.data null: only L is
// written (the caller re-forms the links from L).  reH is linear, so the zero-mode
// generator reH(log U - mean) of nspt.cc:638 is L - mean(L).
// eta: [idx*NE + e] (may be null).
template <int NC, int N>
__device__ __forceinline__ void k_update_body(int blk_, int nblk_, const Geo& g, SerField D, const double2* eta,
                                                                  double stau, SerField Ubase, SerField Uout,
                                                                  SerField L, double2* mm_part) {
  constexpr int NE = NC * NC;
  using LN = Lane<NC, N>;
  __shared__ double2 lds[LN::LDS];
  LN ln(lds);
  const int nrt = g.nrt;
  const int n = min(N, nrt);
  const LinkBlock lb = link_block_at<LN::LPB>(g, block_order(blk_, nblk_));
  const int mu = lb.mu;
  const int site0 = lb.site0;
  const int nvalid = min(LN::LPB, lb.count - site0);
  const int site = link_site(g, min(site0 + ln.j, lb.count - 1));
  const bool ok = ln.lds_ok && (site0 + ln.j) < lb.count;
  const long idx = (long)mu * g.V + site;

  // generator W = D - stau*eta (order 1) into slot A, Ubase into slot B
  // (above 20 orders of SU(3) Ubase is staged from memory after the series)
  constexpr bool HOLD = exp_times_holds_b<NC, N>();
  double2 S[N], W[N];
  gload<NC, N>(D, idx, ln.e, nrt, ok, W);
  const double2 ubf = gflag(Ubase, idx, ok);
  if (eta) W[0] = c_sub(W[0], c_scale(stau, eta[idx * NE + ln.e]));
  // Unew = exp(W) * Ubase, in slot B; Ubase is loaded once W is staged
  if constexpr (HOLD)
    series_exp_times<NC, N>(ln, n, W, ubf, S, [&]() { gload<NC, N>(Ubase, idx, ln.e, nrt, ok, S); });
  else
    ps_exp_times_staged<NC, N>(ln, n, W, ubf, [&]() { stage_global<NC, N>(ln, ln.slotB, Ubase, idx, nrt); });
  unstage(ln, ln.slotB, S);
  if (Uout.data) {
    gstore<NC, N>(Uout, idx, ln.e, nrt, ok, S);
    if (ok && ln.e == 0) Uout.flag[idx] = ubf;
  }
  if (!L.data) return;
  // L = reH(log(Unew)): generator Unew - 1 is already in slot B
  series_fn<NC, N, false>(ln, n, S);
#pragma unroll
  for (int k = 0; k < N; k++) S[k] = reh_elem(ln, ln.slotA, k);
  gstore<NC, N>(L, idx, ln.e, nrt, ok, S);
  if (ok && ln.e == 0 && L.flag) L.flag[idx] = zero2();
  if (!mm_part) return;
  ln.stage(ln.slotB, S);
  __syncthreads();
  for (int o = threadIdx.x; o < nrt * NE; o += 64)
    mm_part[(long)lb.part * nrt * NE + o] = slot_sum<NC, N>(lds + LN::LPB * LN::SLOT, nvalid, o / NE, o % NE);
}

template <int NC, int N>
__global__ __launch_bounds__(64, NSPT_WAVES_PER_EU) void k_update(Geo g, SerField D, const double2* eta,
                                                                  double stau, SerField Ubase, SerField Uout,
                                                                  SerField L, double2* mm_part) {
  k_update_body<NC, N>(blockIdx.x, gridDim.x, g, D, eta, stau, Ubase, Uout, L, mm_part);
}
template <int NC, int N>
__global__ __launch_bounds__(64, NSPT_WAVES_PER_EU) void k_update_tab(const UpdateArgs* __restrict__ t) {
  const UpdateArgs& a = t[blockIdx.y];
  if ((int)blockIdx.x >= a.nblk) return;
  k_update_body<NC, N>(blockIdx.x, a.nblk, a.g, a.D, a.eta, a.stau, a.Ubase, a.Uout, a.L, a.mm_part);
}

// ------------------------------------------------------------ log field
// L = log(U) for every link (FA gauge fixing without zero-mode removal).
template <int NC, int N>
__global__ __launch_bounds__(64, NSPT_WAVES_PER_EU) void k_log(Geo g, SerField U, SerField L) {
  using LN = Lane<NC, N>;
  __shared__ double2 lds[LN::LDS];
  LN ln(lds);
  const int nrt = g.nrt;
  const LinkBlock lb = link_block<LN::LPB>(g);
  const int mu = lb.mu;
  const int site0 = lb.site0;
  const int site = min(site0 + ln.j, g.V - 1);
  const bool ok = ln.lds_ok && (site0 + ln.j) < g.V;
  const long idx = (long)mu * g.V + site;
  double2 S[N];
  if constexpr (use_ps<N>()) {
    gload<NC, N>(U, idx, ln.e, nrt, ok, S);
    ps_series<NC, N, false>(ln, min(N, nrt), S);
  } else {
    stage_global(ln, ln.slotB, U, idx, nrt);
    horner<NC, N, false>(ln, min(N, nrt));
  }
  unstage(ln, ln.slotA, S);
  gstore<NC, N>(L, idx, ln.e, nrt, ok, S);
  if (ok && ln.e == 0 && L.flag) L.flag[idx] = zero2();
}

// ------------------------------------------------------------ zero modes
// X = reH(L - MM_mu), U = exp(X)  (nspt.cc:638-641).  MM: [(mu*nrt + k)*NE + e],
// already multiplied by 1/V.  X is kept: it is log(U) for an FA step.
template <int NC, int N>
__device__ __forceinline__ void k_zero_mode_body(int blk_, int nblk_, const Geo& g, SerField L, const double2* MM,
                                                                     SerField U, SerField X) {
  constexpr int NE = NC * NC;
  using LN = Lane<NC, N>;
  __shared__ double2 lds[LN::LDS];
  LN ln(lds);
  const int nrt = g.nrt;
  const LinkBlock lb = link_block_at<LN::LPB>(g, block_order(blk_, nblk_));
  const int mu = lb.mu;
  const int site0 = lb.site0;
  const int site = min(site0 + ln.j, g.V - 1);
  const bool ok = ln.lds_ok && (site0 + ln.j) < g.V;
  const long idx = (long)mu * g.V + site;
  double2 P[N];
  gload<NC, N>(L, idx, ln.e, nrt, ok, P);
#pragma unroll
  for (int k = 0; k < N; k++)
    if (k < nrt) P[k] = c_sub(P[k], MM[((long)mu * nrt + k) * NE + ln.e]);
  ln.stage(ln.slotA, P);
  __syncthreads();
#pragma unroll
  for (int k = 0; k < N; k++) P[k] = reh_elem(ln, ln.slotA, k);
  gstore<NC, N>(X, idx, ln.e, nrt, ok, P);
  ln.stage(ln.slotB, P);
  series_fn<NC, N, true>(ln, min(N, nrt), P);
  unstage(ln, ln.slotA, P);
  gstore<NC, N>(U, idx, ln.e, nrt, ok, P);
  if (ok && ln.e == 0) U.flag[idx] = one2();
}

template <int NC, int N>
__global__ __launch_bounds__(64, NSPT_WAVES_PER_EU) void k_zero_mode(Geo g, SerField L, const double2* MM,
                                                                     SerField U, SerField X) {
  k_zero_mode_body<NC, N>(blockIdx.x, gridDim.x, g, L, MM, U, X);
}
template <int NC, int N>
__global__ __launch_bounds__(64, NSPT_WAVES_PER_EU) void k_zero_mode_tab(const ZeroModeArgs* __restrict__ t) {
  const ZeroModeArgs& a = t[blockIdx.y];
  if ((int)blockIdx.x >= a.nblk) return;
  k_zero_mode_body<NC, N>(blockIdx.x, a.nblk, a.g, a.L, a.MM, a.U, a.X);
}

// ------------------------------------------------------------ FA: G = exp(reH(W))
// W per site from the transformed buffer fb[(k*NE + e)*V + site] (already
// multiplied by -alpha/(p^2 V)), G = exp(reH(W)) (nspt.cc:902-906).
template <int NC, int N>
__device__ __forceinline__ void k_fa_gexp_body(int blk_, int nblk_, const Geo& g, const double2* fb, SerField G) {
  constexpr int NE = NC * NC;
  using LN = Lane<NC, N>;
  __shared__ double2 lds[LN::LDS];
  LN ln(lds);
  const int nrt = g.nrt;
  const int site0 = block_order(blk_, nblk_) * LN::LPB;
  const int site = min(site0 + ln.j, g.V - 1);
  const bool ok = ln.lds_ok && (site0 + ln.j) < g.V;
  double2 P[N];
#pragma unroll
  for (int k = 0; k < N; k++) P[k] = (k < nrt) ? fb[((long)k * NE + ln.e) * g.V + site] : zero2();
  ln.stage(ln.slotA, P);
  __syncthreads();
#pragma unroll
  for (int k = 0; k < N; k++) P[k] = reh_elem(ln, ln.slotA, k);
  ln.stage(ln.slotB, P);
  series_fn<NC, N, true>(ln, min(N, nrt), P);
  unstage(ln, ln.slotA, P);
  gstore<NC, N>(G, site, ln.e, nrt, ok, P);
  if (ok && ln.e == 0) G.flag[site] = one2();
}

template <int NC, int N>
__global__ __launch_bounds__(64, NSPT_WAVES_PER_EU) void k_fa_gexp(Geo g, const double2* fb, SerField G) {
  k_fa_gexp_body<NC, N>(blockIdx.x, gridDim.x, g, fb, G);
}
template <int NC, int N>
__global__ __launch_bounds__(64, NSPT_WAVES_PER_EU) void k_fa_gexp_tab(const GexpArgs* __restrict__ t) {
  const GexpArgs& a = t[blockIdx.y];
  if ((int)blockIdx.x >= a.nblk) return;
  k_fa_gexp_body<NC, N>(blockIdx.x, a.nblk, a.g, a.fb, a.G);
}

// ------------------------------------------------------------ gauge transform
// U_mu(x) <- G(x) U_mu(x) G(x+mu)^dag, in place (nspt.cc:912-916).  With
// Xexp.data set, U_mu(x) is not read but formed as exp(Xexp - MM_mu) (MM may
// be null): the field after zero-mode removal (nspt.cc:638-641), from the
// projected log L of the update and the mean MM (layout as in k_zero_mode).
#ifndef NSPT_PS_TRANSFORM_PREFETCH
#define NSPT_PS_TRANSFORM_PREFETCH 0
#endif
#ifndef NSPT_PS_TRANSFORM_CHUNK
#define NSPT_PS_TRANSFORM_CHUNK NSPT_STAGE_CHUNK
#endif
template <int NC, int N>
__device__ __forceinline__ void k_gauge_transform_body(int blk_, int nblk_, const Geo& g, SerField G, SerField U,
                                                                           SerField Xexp, const double2* MM) {
  using LN = Lane<NC, N>;
  __shared__ double2 lds[LN::LDS];
  LN ln(lds);
  const int nrt = g.nrt;
  const int M = min(N, nrt) - 1;
  const LinkBlock lb = link_block_at<LN::LPB>(g, block_order(blk_, nblk_));
  const int mu = lb.mu;
  const int site0 = lb.site0;
  const int site = min(site0 + ln.j, g.V - 1);
  const bool ok = ln.lds_ok && (site0 + ln.j) < g.V;
  const long idx = (long)mu * g.V + site;
  const int xp = site_shift(g, site, mu, 1);
  double2 P[N];
  if (Xexp.data) {
    // T = exp(X - MM) G(x+mu)^dag by Horner from the right (the exponential
    // is never formed alone), then U = G(x) T
    // (above 20 orders of SU(3) G(x+mu)^dag is staged from memory after the
    // series instead of held in Q)
    constexpr bool HOLD = exp_times_holds_b<NC, N>();
    constexpr bool PS = use_ps<N>();
    double2 Q[HOLD ? N : 1];
    gload<NC, N>(Xexp, idx, ln.e, nrt, ok, P);
    // the Horner form holds this lane's element of G(x+mu)^dag, staged
    // first; the Paterson-Stockmeyer form holds G(x+mu) itself, loaded once
    // the generator is staged (while the first product runs) and staged
    // daggered after the series
    if constexpr (HOLD && !PS) gload<NC, N>(G, xp, ln.e, nrt, ok, Q);
    const double2 gpf = c_conj(gflag(G, xp, ok));
    if (MM) {
#pragma unroll
      for (int k = 0; k < N; k++)
        if (k < nrt) P[k] = c_sub(P[k], MM[((long)mu * nrt + k) * (NC * NC) + ln.e]);
    }
    if constexpr (HOLD && !PS) {
      ln.stage(ln.slotA, P);
      ln.stage(ln.slotB, Q, true);
      __syncthreads();
      unstage(ln, ln.slotB, Q);  // this lane's element of G(x+mu)^dag
    }
    // G(x) is loaded into the free registers while the exponential runs
    // (above 24 orders those registers are not free: it is loaded after;
    // the Paterson-Stockmeyer form holds its generator in them)
    constexpr bool PF = N <= 24 && (!PS || NSPT_PS_TRANSFORM_PREFETCH);
    const double2 gf = gflag(G, site, ok);
    double2 Gx[PS && PF ? N : 1];
    if constexpr (PS) {
      if constexpr (PF && !HOLD) gload<NC, N>(G, site, ln.e, nrt, ok, Gx);
      if constexpr (HOLD)
        ps_exp_times<NC, N>(ln, M + 1, P, gpf, Q, [&]() {
          gload<NC, N>(G, xp, ln.e, nrt, ok, Q);
          if constexpr (PF) gload<NC, N>(G, site, ln.e, nrt, ok, Gx);
        }, true);
      else
        ps_exp_times_staged<NC, N>(ln, M + 1, P, gpf,
                                   [&]() { stage_global<NC, N>(ln, ln.slotB, G, xp, nrt, true); });
    } else {
      if constexpr (PF) gload<NC, N>(G, site, ln.e, nrt, ok, P);
      horner_exp_times<NC, N>(ln, M + 1, gpf, Q);
    }
    if constexpr (PF && PS)
      ln.stage(ln.slotA, Gx);
    else if constexpr (PF)
      ln.stage(ln.slotA, P);
    else
      stage_global<NC, N, PS ? NSPT_PS_TRANSFORM_CHUNK : 1>(ln, ln.slotA, G, site, nrt);
    __syncthreads();
    // (the full-length products of the Paterson-Stockmeyer instances take its
    // larger order tiles)
    const double2 nf = mul_lds<NC, N, PS ? NSPT_QT_PS : NSPT_QT>(ln, gf, gpf, true, M, ln.slotB, false);
    __syncthreads();
    unstage(ln, ln.slotB, P);
    gstore<NC, N>(U, idx, ln.e, nrt, ok, P);
    if (ok && ln.e == 0) U.flag[idx] = nf;
    return;
  }
  stage_global(ln, ln.slotA, G, site, nrt);
  stage_global(ln, ln.slotB, U, idx, nrt);
  const double2 uf = gflag(U, idx, ok);
  const double2 gf = gflag(G, site, ok);
  // G(x+mu) is loaded while T = G(x) U_mu(x) is formed in place in slot A
  gload<NC, N>(G, xp, ln.e, nrt, ok, P);
  const double2 gpf = c_conj(gflag(G, xp, ok));
  __syncthreads();
  const double2 tf = mul_lds<NC, N>(ln, gf, uf, true, M, ln.slotA, false);
  __syncthreads();
  ln.stage(ln.slotB, P, true);
  __syncthreads();
  const double2 nf = mul_lds<NC, N>(ln, tf, gpf, true, M, ln.slotB, false);
  __syncthreads();
  unstage(ln, ln.slotB, P);
  gstore<NC, N>(U, idx, ln.e, nrt, ok, P);
  if (ok && ln.e == 0) U.flag[idx] = nf;
}

template <int NC, int N>
__global__ __launch_bounds__(64, NSPT_WAVES_PER_EU) void k_gauge_transform(Geo g, SerField G, SerField U,
                                                                           SerField Xexp, const double2* MM) {
  k_gauge_transform_body<NC, N>(blockIdx.x, gridDim.x, g, G, U, Xexp, MM);
}
template <int NC, int N>
__global__ __launch_bounds__(64, NSPT_WAVES_PER_EU) void k_gauge_transform_tab(const TransformArgs* __restrict__ t) {
  const TransformArgs& a = t[blockIdx.y];
  if ((int)blockIdx.x >= a.nblk) return;
  k_gauge_transform_body<NC, N>(blockIdx.x, a.nblk, a.g, a.G, a.U, a.Xexp, a.MM);
}

// ------------------------------------------------------------ checkerboard gauge fixing
// One pass of the sequential gauge fixing of nspt.cc:660-862 over the sites
// of one parity (g.sub = 1 + 4 parity), which share no link: per site
// W = sum_mu U_mu(x) - U_mu(x-mu), W <- Trless(W - W^dag), G = exp(alpha W),
// then U_mu(x) <- G U_mu(x) and U_mu(x-mu) <- U_mu(x-mu) G^dag
// (exp(-alpha W) = G^dag, W antihermitian).
template <int NC, int N>
__global__ __launch_bounds__(64, NSPT_WAVES_PER_EU) void k_sgf(Geo g, SerField U, double alpha) {
  constexpr int NE = NC * NC;
  using LN = Lane<NC, N>;
  __shared__ double2 lds[LN::LDS];
  LN ln(lds);
  const int nrt = g.nrt;
  const int M = min(N, nrt) - 1;
  const int cnt = g.V / 2;
  const int k0 = block_order() * LN::LPB;
  const int site = link_site(g, min(k0 + ln.j, cnt - 1));
  const bool ok = ln.lds_ok && (k0 + ln.j) < cnt;
  const long V = g.V;
  double2 W[N], P[N];
#pragma unroll
  for (int k = 0; k < N; k++) W[k] = zero2();
  for (int mu = 0; mu < 4; mu++) {
    gload<NC, N>(U, mu * V + site, ln.e, nrt, ok, P);
#pragma unroll
    for (int k = 0; k < N; k++) W[k] = c_add(W[k], P[k]);
    int cm = TW0;
    const int xm = site_move_tw(g, site, mu, -1, cm);
    gload_tw<NC, N>(g, ln, U, mu * V + xm, cm, nrt, ok, P);
#pragma unroll
    for (int k = 0; k < N; k++) W[k] = c_sub(W[k], P[k]);
  }
  // W - W^dag: this lane's element minus the conjugate of the transposed one
  ln.stage(ln.slotA, W);
  __syncthreads();
#pragma unroll
  for (int k = 0; k < N; k++) W[k] = c_sub(ln.slotA[k * NE + ln.e], c_conj(ln.slotA[k * NE + ln.c * NC + ln.r]));
  __syncthreads();
  // Trless (QCDpt.cc:124-136): Tr/NC off the diagonal of every order
  ln.stage(ln.slotA, W);
  __syncthreads();
  if (ln.r == ln.c) {
#pragma unroll
    for (int k = 0; k < N; k++) {
      double2 tr = zero2();
#pragma unroll
      for (int q = 0; q < NC; q++) tr = c_add(tr, ln.slotA[k * NE + q * NC + q]);
      W[k] = c_sub(W[k], c_scale(1.0 / NC, tr));
    }
  }
  __syncthreads();
  // G = exp(alpha W)
#pragma unroll
  for (int k = 0; k < N; k++) W[k] = c_scale(alpha, W[k]);
  ln.stage(ln.slotB, W);
  series_fn<NC, N, true>(ln, M + 1, W);
  unstage(ln, ln.slotA, W);  // W holds G from here on
  __syncthreads();
  for (int mu = 0; mu < 4; mu++) {
    // U_mu(x) <- G U_mu(x), in place in slot B
    const long i1 = mu * V + site;
    ln.stage(ln.slotA, W);
    stage_global(ln, ln.slotB, U, i1, nrt);
    const double2 f1 = gflag(U, i1, ok);
    __syncthreads();
    const double2 n1 = mul_lds<NC, N>(ln, one2(), f1, true, M, ln.slotB, false);
    __syncthreads();
    unstage(ln, ln.slotB, P);
    gstore<NC, N>(U, i1, ln.e, nrt, ok, P);
    if (ok && ln.e == 0) U.flag[i1] = n1;
    __syncthreads();
    // U_mu(x-mu) <- U_mu(x-mu) G^dag, in place in slot A (on a twisted
    // lattice: the twisted copy times G^dag, stored back untwisted)
    int c2 = TW0;
    const long i2 = mu * V + site_move_tw(g, site, mu, -1, c2);
    stage_global_tw(g, ln, ln.slotA, U, i2, c2, nrt);
    ln.stage(ln.slotB, W, true);
    const double2 f2 = gflag(U, i2, ok);
    __syncthreads();
    const double2 n2 = mul_lds<NC, N>(ln, f2, one2(), true, M, ln.slotA, false);
    __syncthreads();
    unstage(ln, ln.slotA, P);
    gstore_tw<NC, N>(g, ln, U, i2, c2, nrt, ok, P);
    if (ok && ln.e == 0) U.flag[i2] = n2;
    __syncthreads();
  }
}

// ------------------------------------------------------------ rectangle staples
// R_mu(x) = sum_{nu != mu} staple2x1(x, mu, nu) (QCDenvNODEpt_test.h:579-611):
// per plane the six 5-link staples of the 2x1 and 1x2 rectangles through
// the link.  Every rectangle has a long side of two collinear links outside
// the link itself, so with the 2-link field L2_d(x) = U_d(x) U_d(x+d) each
// staple is a path of 3 or 4 factors; and two shapes of every plane start
// with U_mu(x+mu), two end with U_mu(x-mu), so those are summed over the
// planes first:
//   R = U_mu(x+mu) sum_nu (B1 + B4) + sum_nu (A2 + A5) U_mu(x-mu)
//       + sum_nu (S0 + S3),
//   B1 = U_nu(x+2mu) L2_mu(x+nu)^+ U_nu(x)^+,
//   B4 = U_nu(x+2mu-nu)^+ L2_mu(x-nu)^+ U_nu(x-nu),
//   A2 = U_nu(x+mu) L2_mu(x-mu+nu)^+ U_nu(x-mu)^+,
//   A5 = U_nu(x+mu-nu)^+ L2_mu(x-mu-nu)^+ U_nu(x-mu-nu),
//   S0 = L2_nu(x+mu) U_mu(x+2nu)^+ L2_nu(x)^+,
//   S3 = L2_nu(x+mu-2nu)^+ U_mu(x-2nu)^+ L2_nu(x-2nu):
// 38 products per link (72 as the reference forms them).  Paths are formed
// one factor at a time in place in slot A and summed into Acc (registers);
// the three partial sums of R go through the R field itself.  With obs2: per
// block, Tr(U_mu(x) R_mu(x)) per order (the rectangle plaquette of
// nspt.cc:406-410), as the drift's plaquette rows.
template <int NC, int N>
__global__ __launch_bounds__(64, NSPT_WAVES_PER_EU) void k_rect(Geo g, SerField U, SerField L2, SerField R,
                                                                double* obs2) {
  constexpr int NE = NC * NC;
  using LN = Lane<NC, N>;
  __shared__ double2 lds[LN::LDS];
  LN ln(lds);
  const int nrt = g.nrt;
  const int M = min(N, nrt) - 1;
  const LinkBlock lb = link_block<LN::LPB>(g);
  const int mu = lb.mu;
  const int site0 = lb.site0;
  const int nvalid = min(LN::LPB, lb.count - site0);
  const int site = link_site(g, min(site0 + ln.j, lb.count - 1));
  const bool ok = ln.lds_ok && (site0 + ln.j) < lb.count;
  const long V = g.V;
  const long idx = mu * V + site;
  // A step stages one factor: offset (dm along mu, dn along nu, plane
  // p: nu = the p-th direction other than mu), direction (0 mu, 1 nu),
  // dagger, source (0 U, 1 L2, 2 the sum Acc), position in its path (0
  // first, 1 middle, 2 last), and what follows a last factor (0 keep
  // summing into Acc, 1 R = Acc, 2 R += Acc).  Phase B (steps 0-17), then
  // U_mu(x+mu) and Acc (18, 19), phase A (20-37), Acc and U_mu(x-mu) (38,
  // 39), phase S (40-57); a phase takes 6 factors per plane.
  constexpr int NSTEP = 58;
  struct Step {
    signed char dm, dn, dir, dag, src, pos, post, p;
  };
  static constexpr signed char PH[3][6][5] = {
      {{2, 0, 1, 0, 0}, {0, 1, 0, 1, 1}, {0, 0, 1, 1, 0}, {2, -1, 1, 1, 0}, {0, -1, 0, 1, 1}, {0, -1, 1, 0, 0}},
      {{1, 0, 1, 0, 0}, {-1, 1, 0, 1, 1}, {-1, 0, 1, 1, 0}, {1, -1, 1, 1, 0}, {-1, -1, 0, 1, 1}, {-1, -1, 1, 0, 0}},
      {{1, 0, 1, 0, 1}, {0, 2, 0, 1, 0}, {0, 0, 1, 1, 1}, {1, -2, 1, 1, 1}, {0, -2, 0, 1, 0}, {0, -2, 1, 0, 1}}};
  struct Table {
    Step s[NSTEP];
  };
  // the step list, built at compile time and packed into one word per step
  static constexpr auto table = [] {
    Table t{};
    int n = 0;
    for (int ph = 0; ph < 3; ph++) {
      for (int r = 0; r < 18; r++) {
        const signed char* e = PH[ph][r % 6];
        const int k = r % 6;
        t.s[n++] = {e[0], e[1], e[2], e[3], e[4], (signed char)(k % 3), (signed char)(ph == 2 && r == 17 ? 2 : 0),
                    (signed char)(r / 6)};
      }
      if (ph == 0) {
        t.s[n++] = {1, 0, 0, 0, 0, 0, 0, 0};  // U_mu(x+mu)
        t.s[n++] = {0, 0, 0, 0, 2, 2, 1, 0};  // times sum_nu (B1 + B4) -> R
      } else if (ph == 1) {
        t.s[n++] = {0, 0, 0, 0, 2, 0, 0, 0};   // sum_nu (A2 + A5)
        t.s[n++] = {-1, 0, 0, 0, 0, 2, 2, 0};  // times U_mu(x-mu) -> R +=
      }
    }
    return t;
  }();
  struct Fac {
    long li;
    int cl, pos, src, post;
    bool dg;
  };
  auto fac_of = [&](int step) {
    const Step st = table.s[uniform(step)];
    Fac f;
    f.cl = TW0;
    const int nu = (st.p < mu) ? st.p : st.p + 1;
    const int x = site_move_tw(g, site_move_tw(g, site, mu, st.dm, f.cl), nu, st.dn, f.cl);
    f.li = (long)(st.dir ? nu : mu) * V + x;
    f.dg = st.dag != 0;
    f.src = st.src;
    f.pos = st.pos;
    f.post = st.post;
    return f;
  };
  // the field of a factor (pointers picked one by one: a runtime pick of a
  // whole struct goes through scratch)
  auto field = [&](int src) {
    SerField f;
    f.data = src == 1 ? L2.data : U.data;
    f.flag = src == 1 ? L2.flag : U.flag;
    f.P = U.P;
    return f;
  };
  double2 Acc[N];
#pragma unroll
  for (int k = 0; k < N; k++) Acc[k] = zero2();
  double2 af = zero2(), rf = zero2(), tf = zero2();
  // the next factor is loaded into registers while the current product runs
  constexpr bool PF = N <= 24;
  constexpr int NP = PF ? N : 1;
  double2 PB[NP];
  Fac cur = fac_of(0);
  if constexpr (PF) gload_tw<NC, N>(g, ln, field(cur.src), cur.li, cur.cl, nrt, ok, PB);
  for (int step = 0; step < NSTEP; step++) {
    asm volatile("" ::: "memory");
    double2* slot = (cur.pos == 0) ? ln.slotA : ln.slotB;
    double2 fl;
    if (cur.src == 2) {
      // the sum Acc as a factor; Acc restarts from zero
      ln.stage(slot, Acc);
      fl = af;
#pragma unroll
      for (int k = 0; k < N; k++) Acc[k] = zero2();
      af = zero2();
    } else {
      const SerField fc = field(cur.src);
      fl = gflag(fc, cur.li, ok);
      if (cur.dg) fl = c_conj(fl);
      if constexpr (PF)
        ln.stage(slot, PB, cur.dg);
      else
        stage_global_tw(g, ln, slot, fc, cur.li, cur.cl, nrt, cur.dg);
    }
    const int pos = cur.pos, post = cur.post;
    if (step + 1 < NSTEP) {
      cur = fac_of(step + 1);
      if constexpr (PF)
        if (cur.src != 2) gload_tw<NC, N>(g, ln, field(cur.src), cur.li, cur.cl, nrt, ok, PB);
    }
    if (pos == 0) {
      tf = fl;
      continue;
    }
    __syncthreads();
    if (pos == 1) {
      tf = mul_lds<NC, N>(ln, tf, fl, true, M, ln.slotA, false);
    } else {
      af = c_add(af, mul<NC, N>(ln, tf, fl, true, Acc, M, false, nullptr, false, 1.0));
    }
    __syncthreads();
    if (post) {
      // R = Acc or R += Acc; Acc restarts from zero
      gaccum<NC, N>(R, idx, ln.e, nrt, ok, Acc, post == 1);
      rf = post == 1 ? af : c_add(rf, af);
#pragma unroll
      for (int k = 0; k < N; k++) Acc[k] = zero2();
      af = zero2();
    }
  }
  if (ok && ln.e == 0) R.flag[idx] = rf;
  if (!obs2) return;
  // Tr(U_mu(x) R_mu(x)) per order, summed over the block's links
  __shared__ double2 flagbuf[LN::LPB];
  stage_global(ln, ln.slotA, U, idx, nrt);
  stage_global(ln, ln.slotB, R, idx, nrt);
  const double2 uf = gflag(U, idx, ok);
  __syncthreads();
  const double2 wf = mul_lds<NC, N>(ln, uf, rf, true, M, ln.slotB, false);
  if (ln.lds_ok && ln.e == 0) flagbuf[ln.j] = wf;
  __syncthreads();
  const int lane = threadIdx.x;
  if (lane <= nrt) {
    double2 acc = zero2();
    for (int jj = 0; jj < nvalid; jj++) {
      if (lane == 0) {
        acc = c_add(acc, c_scale((double)NC, flagbuf[jj]));
      } else {
        const double2* m = lds + LN::LPB * LN::SLOT + jj * LN::SLOT + (lane - 1) * NE;
#pragma unroll
        for (int q = 0; q < NC; q++) acc = c_add(acc, m[q * NC + q]);
      }
    }
    double* op = obs2 + (long)lb.part * 2 * (nrt + 1);
    op[2 * lane] = acc.x;
    op[2 * lane + 1] = acc.y;
  }
}

// ------------------------------------------------------------ Wilson loops
// Long links out_mu(x) = A_mu(x) A'_mu(x + k mu) with A' = A: the 2-link
// field (k = 1, A = U) and the 4-link field (k = 2, A = the 2-link field)
// of the loop measurement.
// Link blocks per workgroup (above 24 orders the registers of a prefetch
// are not free: one block, as many workgroups as blocks).
template <int N>
constexpr int link_pair_blocks() {
  return N <= 24 ? 4 : 1;
}
template <int NC, int N>
__global__ __launch_bounds__(64, NSPT_WAVES_PER_EU) void k_link_pair(Geo g, SerField A, SerField out, int k,
                                                                     int nblk) {
  using LN = Lane<NC, N>;
  __shared__ double2 lds[LN::LDS];
  LN ln(lds);
  const int nrt = g.nrt;
  const int M = min(N, nrt) - 1;
  // NB link blocks per workgroup, one after the other; the next block's two
  // links are loaded while the current product runs (a single product per
  // block leaves the loads' latency exposed otherwise)
  constexpr int NB = link_pair_blocks<N>();
  const int b0 = block_order() * NB;
  double2 PA[N], PB[N], P[N];
  double2 fa, fb;
  auto load = [&](int b) {
    const LinkBlock lb = link_block_at<LN::LPB>(g, b);
    const int site = min(lb.site0 + ln.j, g.V - 1);
    const long ix = (long)lb.mu * g.V + site;
    int c2 = TW0;
    const long ix2 = (long)lb.mu * g.V + site_shift_far_tw(g, site, lb.mu, k, c2);
    const bool okb = ln.lds_ok && (lb.site0 + ln.j) < g.V;
    gload<NC, N>(A, ix, ln.e, nrt, okb, PA);
    gload_tw<NC, N>(g, ln, A, ix2, c2, nrt, okb, PB);
    fa = gflag_all(A, ix);
    fb = gflag_all(A, ix2);
  };
  load(b0);
  for (int i = 0; i < NB && b0 + i < nblk; i++) {
    const LinkBlock lb = link_block_at<LN::LPB>(g, b0 + i);
    const int site = min(lb.site0 + ln.j, g.V - 1);
    const bool ok = ln.lds_ok && (lb.site0 + ln.j) < g.V;
    const long idx = (long)lb.mu * g.V + site;
    const double2 a0 = fa, bf = fb;
    ln.stage(ln.slotA, PA);
    ln.stage(ln.slotB, PB);
    __syncthreads();
    if (i + 1 < NB && b0 + i + 1 < nblk) load(b0 + i + 1);
    const double2 f = mul_lds<NC, N>(ln, a0, bf, true, M, ln.slotB, false);
    __syncthreads();
    unstage(ln, ln.slotB, P);
    gstore<NC, N>(out, idx, ln.e, nrt, ok, P);
    if (ok && ln.e == 0) out.flag[idx] = f;
    __syncthreads();  // the slots are restaged for the next block
  }
}

// One rectangle type of one plane per block (ComputeLoops, nspt.cc:1156-1312):
// the lm x ln loop at x in plane (mu, nu) is Tr(T1 T2^dag) with
//   T1 = F(lm)_mu(x) F(ln)_nu(x + lm mu),  T2 = F(ln)_nu(x) F(lm)_mu(x + ln nu),
// F(1) = U, F(2) = the 2-link, F(4) = the 4-link field.  Rectangles
// r = 0..5: 2x1, 1x2, 2x2, 4x2, 2x4, 4x4; planes p = 0..5: (0,1), (0,2),
// (0,3), (1,2), (1,3), (2,3).  Block b -> site block b % nb, plane
// (b / nb) % 6, rectangle b / (6 nb); its sum over the block's sites of Tr,
// orders 0..nrt (complex), goes to row b of part (2 (nrt+1) doubles).
template <int NC, int N>
__global__ __launch_bounds__(64, NSPT_WAVES_PER_EU) void k_loops(Geo g, SerField U, SerField L2, SerField L4,
                                                                 double* part) {
  constexpr int NE = NC * NC;
  using LN = Lane<NC, N>;
  __shared__ double2 lds[LN::LDS];
  static_assert((N + 1) * 64 <= LN::LDS, "lane sums fit in the slots");
  LN ln(lds);
  const int nrt = g.nrt;
  const int M = min(N, nrt) - 1;
  const int nb = g.nb_mu;
  const int b = block_order();
  const int sb = b % nb, plane = (b / nb) % 6, rect = b / (6 * nb);
  const int mu = plane < 3 ? 0 : (plane < 5 ? 1 : 2);
  const int nu = plane < 3 ? plane + 1 : (plane < 5 ? plane - 1 : 3);
  const int lm = (rect == 1) ? 1 : (rect == 3 || rect == 5) ? 4 : 2;
  const int lnu = (rect == 0) ? 1 : (rect == 4 || rect == 5) ? 4 : 2;
  // (pointers picked one by one: a runtime pick of a whole struct goes through scratch)
  auto F = [&](int len) {
    SerField f;
    f.data = len == 1 ? U.data : (len == 2 ? L2.data : L4.data);
    f.flag = len == 1 ? U.flag : (len == 2 ? L2.flag : L4.flag);
    f.P = U.P;
    return f;
  };
  const int site0 = sb * LN::LPB;
  const int site = min(site0 + ln.j, g.V - 1);
  const bool ok = ln.lds_ok && (site0 + ln.j) < g.V;
  const long V = g.V;
  int c2 = TW0, c4 = TW0;
  const long i1 = mu * V + site, i2 = nu * V + site_shift_far_tw(g, site, mu, lm, c2);
  const long i3 = nu * V + site, i4 = mu * V + site_shift_far_tw(g, site, nu, lnu, c4);
  double2 P[N], Q[N];
  // T1
  stage_global(ln, ln.slotA, F(lm), i1, nrt);
  stage_global_tw(g, ln, ln.slotB, F(lnu), i2, c2, nrt);
  const double2 fa = gflag(F(lm), i1, ok), fb = gflag(F(lnu), i2, ok);
  __syncthreads();
  const double2 f1 = mul_lds<NC, N>(ln, fa, fb, true, M, ln.slotA, false);
  __syncthreads();
  unstage(ln, ln.slotA, P);
  __syncthreads();
  // T2
  stage_global(ln, ln.slotA, F(lnu), i3, nrt);
  stage_global_tw(g, ln, ln.slotB, F(lm), i4, c4, nrt);
  const double2 fc = gflag(F(lnu), i3, ok), fd = gflag(F(lm), i4, ok);
  __syncthreads();
  const double2 f2 = mul_lds<NC, N>(ln, fc, fd, true, M, ln.slotA, false);
  __syncthreads();
  unstage(ln, ln.slotA, Q);
  // this lane's share of Tr(T1 T2^dag), order k+1 in acc[k]: the pairs
  // P_i conj(Q_j) with i + j + 1 = k over element (r, c), and on the
  // diagonal the flag terms f1 conj(Q_k) + conj(f2) P_k
#pragma unroll
  for (int k = 0; k < N; k++) {
    P[k] = (k <= M) ? P[k] : zero2();
    Q[k] = (k <= M) ? Q[k] : zero2();
  }
  double2 acc[N];
#pragma unroll
  for (int k = 0; k < N; k++) acc[k] = zero2();
#pragma unroll
  for (int i = 0; i < N; i++)
#pragma unroll
    for (int j = 0; i + j + 1 < N; j++) c_mac(acc[i + j + 1], P[i], c_conj(Q[j]));
  if (ln.r == ln.c) {
#pragma unroll
    for (int k = 0; k < N; k++) acc[k] = c_add(acc[k], c_add(c_mul(f1, c_conj(Q[k])), c_mul(c_conj(f2), P[k])));
  }
  const double2 t0 = (ok && ln.e == 0) ? c_scale((double)NC, c_mul(f1, c_conj(f2))) : zero2();
  __syncthreads();
  const int lane = threadIdx.x;
#pragma unroll
  for (int k = 0; k < N; k++) lds[k * 64 + lane] = ok ? acc[k] : zero2();
  lds[N * 64 + lane] = t0;
  __syncthreads();
  if (lane <= nrt) {
    const double2* col = lds + (lane == 0 ? N : lane - 1) * 64;
    double2 sum = zero2();
    for (int l = 0; l < 64; l++) sum = c_add(sum, col[l]);
    double* row = part + (long)b * 2 * (nrt + 1);
    row[2 * lane] = sum.x;
    row[2 * lane + 1] = sum.y;
  }
}

// ------------------------------------------------------------ instance table
template <int NC, int N>
KernelSet make_kernel_set() {
  KernelSet s;
  s.nc = NC;
  s.nmax = N;
  s.big = NSPT_BIG_FIELDS != 0;
  s.lpb = Lane<NC, N>::LPB;
  s.lpw = NSPT_LPW;
  s.wpe = NSPT_WAVES_PER_EU;
  s.qt = NSPT_QT;
  s.drift = [](int nblk, hipStream_t st, Geo g, SerField U, SerField D, SerField D1, double tau_g, double rk2c,
               int mode, double* obs, SerField R, double c1) {
    hipLaunchKernelGGL((k_drift<NC, N>), dim3(nblk), dim3(64), 0, st, g, U, D, D1, tau_g, rk2c, mode, obs, R, c1);
  };
  s.plaq_field = [](int nblk, hipStream_t st, Geo g, SerField U, SerField P) {
    hipLaunchKernelGGL((k_plaq_field<NC, N>), dim3(nblk), dim3(64), 0, st, g, U, P);
  };
  s.drift_p = [](int nblk, hipStream_t st, Geo g, SerField U, SerField P, SerField D, SerField D1, double tau_g,
                 double rk2c, int mode, double* obs, SerField R, double c1) {
    hipLaunchKernelGGL((k_drift_p<NC, N>), dim3(nblk), dim3(64), 0, st, g, U, P, D, D1, tau_g, rk2c, mode, obs, R,
                       c1);
  };
  s.sgf = [](int nblk, hipStream_t st, Geo g, SerField U, double alpha) {
    hipLaunchKernelGGL((k_sgf<NC, N>), dim3(nblk), dim3(64), 0, st, g, U, alpha);
  };
  s.rect = [](int nblk, hipStream_t st, Geo g, SerField U, SerField L2, SerField R, double* obs2) {
    hipLaunchKernelGGL((k_rect<NC, N>), dim3(nblk), dim3(64), 0, st, g, U, L2, R, obs2);
  };
  s.update = [](int nblk, hipStream_t st, Geo g, SerField D, const double2* eta, double stau, SerField Ub,
                SerField Uo, SerField L, double2* mm) {
    hipLaunchKernelGGL((k_update<NC, N>), dim3(nblk), dim3(64), 0, st, g, D, eta, stau, Ub, Uo, L, mm);
  };
  s.logf = [](int nblk, hipStream_t st, Geo g, SerField U, SerField L) {
    hipLaunchKernelGGL((k_log<NC, N>), dim3(nblk), dim3(64), 0, st, g, U, L);
  };
  s.zero_mode = [](int nblk, hipStream_t st, Geo g, SerField L, const double2* MM, SerField U, SerField X) {
    hipLaunchKernelGGL((k_zero_mode<NC, N>), dim3(nblk), dim3(64), 0, st, g, L, MM, U, X);
  };
  s.gexp = [](int nblk, hipStream_t st, Geo g, const double2* fb, SerField G) {
    hipLaunchKernelGGL((k_fa_gexp<NC, N>), dim3(nblk), dim3(64), 0, st, g, fb, G);
  };
  s.transform = [](int nblk, hipStream_t st, Geo g, SerField G, SerField U, SerField Xexp, const double2* MM) {
    hipLaunchKernelGGL((k_gauge_transform<NC, N>), dim3(nblk), dim3(64), 0, st, g, G, U, Xexp, MM);
  };
  // replica-table launches: grid (gx, nrep), gx a multiple of 8
  s.drift_tab = [](int gx, int nrep, hipStream_t st, const DriftArgs* t) {
    hipLaunchKernelGGL((k_drift_tab<NC, N>), dim3(gx, nrep), dim3(64), 0, st, t);
  };
  s.plaq_field_tab = [](int gx, int nrep, hipStream_t st, const PlaqArgs* t) {
    hipLaunchKernelGGL((k_plaq_field_tab<NC, N>), dim3(gx, nrep), dim3(64), 0, st, t);
  };
  s.drift_p_tab = [](int gx, int nrep, hipStream_t st, const DriftPArgs* t) {
    hipLaunchKernelGGL((k_drift_p_tab<NC, N>), dim3(gx, nrep), dim3(64), 0, st, t);
  };
  s.update_tab = [](int gx, int nrep, hipStream_t st, const UpdateArgs* t) {
    hipLaunchKernelGGL((k_update_tab<NC, N>), dim3(gx, nrep), dim3(64), 0, st, t);
  };
  s.zero_mode_tab = [](int gx, int nrep, hipStream_t st, const ZeroModeArgs* t) {
    hipLaunchKernelGGL((k_zero_mode_tab<NC, N>), dim3(gx, nrep), dim3(64), 0, st, t);
  };
  s.gexp_tab = [](int gx, int nrep, hipStream_t st, const GexpArgs* t) {
    hipLaunchKernelGGL((k_fa_gexp_tab<NC, N>), dim3(gx, nrep), dim3(64), 0, st, t);
  };
  s.transform_tab = [](int gx, int nrep, hipStream_t st, const TransformArgs* t) {
    hipLaunchKernelGGL((k_gauge_transform_tab<NC, N>), dim3(gx, nrep), dim3(64), 0, st, t);
  };
  s.link_pair = [](int nblk, hipStream_t st, Geo g, SerField A, SerField out, int k) {
    constexpr int nb = link_pair_blocks<N>();
    hipLaunchKernelGGL((k_link_pair<NC, N>), dim3((nblk + nb - 1) / nb), dim3(64), 0, st, g, A, out, k, nblk);
  };
  s.loops = [](int nblk, hipStream_t st, Geo g, SerField U, SerField L2, SerField L4, double* part) {
    hipLaunchKernelGGL((k_loops<NC, N>), dim3(nblk), dim3(64), 0, st, g, U, L2, L4, part);
  };
  return s;
}

}  // inline namespace NSPT_VARIANT
}  // namespace nspt
