// nspt_adjoint.hip -- the Langevin drift as a reverse-mode gradient of the
// order-truncated Wilson action (no counterpart in the reference, whose
// drift is the staple sum of nspt.cc:383-417).
//
// Field: generalised series, order 0 a full matrix (random configurations
// with Haar-random order 0), orders k = 0..NG-1 stored SoA as
// data[(k*NE + e)*P + idx], idx = mu*V + site.  Action truncated at order K:
//   f = sum_{x, mu<nu} Re Tr P_{mu nu}(x)_K,
//   P = U_mu(x) U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag.
// Gradient convention G = df/dRe U + i df/dIm U (df = Re Tr(G^dag dU)).
//
// k_plaq_adjoint: one workgroup per LPB plaquettes of one plane.  Forward:
// T1 = U1 U2, T2 = T1 U3^dag.  Backward from Pbar = delta_{kK} 1:
//   T2bar = adjA(Pbar, U4^dag)   U4bar = adjB(T2, Pbar)^dag
//   T1bar = adjA(T2bar, U3^dag)  U3bar = adjB(T1, T2bar)^dag
//   U1bar = adjA(T1bar, U2)      U2bar = adjB(U1, T1bar)
// with adjA(Cbar, B)_i = sum_k Cbar_k B_{k-i}^dag and
//      adjB(A, Cbar)_j = sum_k A_{k-j}^dag Cbar_k.
// Both adjoints are ordinary truncated products once the cotangent is
// staged with its orders reversed (order index n of the product is order
// NG-1-n of the adjoint), so they run on the LDS product of
// nspt_device.hpp (full Cauchy form).  Intermediate series are recomputed
// per plaquette and kept in registers; nothing is stored between kernels
// except each link's share of the gradient: a link belongs to six
// plaquettes, each writing one of its six slots (no races), which
// k_adjoint_drift sums in a fixed order.  The slot buffer is a field of
// 6 x 4V series (plane stride 6 x 4V), series index slot*4V + link.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/nspt.h"
#include "nspt_device.hpp"
#include "nspt_internal.hpp"

namespace nspt {
namespace {

__device__ __forceinline__ int shift_site(const Geo& g, int s, int d, int delta) {
  const int x = (s / g.st[d]) % g.L[d];
  int xn = x + delta;
  xn += (xn < 0) ? g.L[d] : 0;
  xn -= (xn >= g.L[d]) ? g.L[d] : 0;
  return s + (xn - x) * g.st[d];
}

// index of direction b among the three directions other than a
__device__ __forceinline__ int other_index(int a, int b) { return b < a ? b : b - 1; }

// A field of ng orders (0..ng-1, ng <= NG at run time) is read into the
// compiled NG orders with zeros above: the products then form the same
// orders 0..ng-1 as an instance compiled for ng (orders of a truncated
// product only feed upward, and the cotangents start at order ng-1), and
// only those are stored.
template <int NC, int NG>
__device__ __forceinline__ void load_ser(const Lane<NC, NG>& ln, const double2* f, long P, long idx, int ng,
                                         double2 (&X)[NG]) {
  constexpr int NE = NC * NC;
#pragma unroll
  for (int k = 0; k < NG; k++) X[k] = (k < ng) ? f[((long)k * NE + ln.e) * P + idx] : zero2();
}
template <int NC, int NG>
__device__ __forceinline__ void store_ser(const Lane<NC, NG>& ln, double2* f, long P, long idx, bool ok, int ng,
                                          const double2 (&X)[NG], bool dag) {
  constexpr int NE = NC * NC;
  // the dagger of a lane-owned series needs the transposed element: go
  // through LDS (the caller has a free slot in ln.slotA)
  if (!dag) {
    if (ok)
#pragma unroll
      for (int k = 0; k < NG; k++)
        if (k < ng) f[((long)k * NE + ln.e) * P + idx] = X[k];
    return;
  }
  __syncthreads();
  ln.stage(ln.slotA, X, true);
  __syncthreads();
  if (ok)
#pragma unroll
    for (int k = 0; k < NG; k++)
      if (k < ng) f[((long)k * NE + ln.e) * P + idx] = ln.slotA[k * NE + ln.e];
}

// A series of ng orders staged straight from memory into an LDS slot (or
// its dagger), zero above ng: loads issued eight orders at a time, so that
// no register array is held for it across a product.
template <int NC, int NG>
__device__ __forceinline__ void stage_ser(const Lane<NC, NG>& ln, double2* slot, const double2* f, long P, long idx,
                                          int ng, bool dag) {
  constexpr int NE = NC * NC;
  if (!ln.lds_ok) return;
  const int pos = dag ? (ln.c * NC + ln.r) : ln.e;
  const double sy = dag ? -1.0 : 1.0;
  constexpr int CH = 8;
#pragma unroll
  for (int k0 = 0; k0 < NG; k0 += CH) {
    double2 v[CH];
#pragma unroll
    for (int i = 0; i < CH; i++)
      if (k0 + i < NG) v[i] = (k0 + i < ng) ? f[((long)(k0 + i) * NE + ln.e) * P + idx] : zero2();
#pragma unroll
    for (int i = 0; i < CH; i++)
      if (k0 + i < NG) slot[(k0 + i) * NE + pos] = make_double2(v[i].x, sy * v[i].y);
  }
}
// The cotangent of the plaquette's order-K trace, delta_{kK} 1, staged with
// its orders reversed (index n holds order NG-1-n).
template <int NC, int NG>
__device__ __forceinline__ void stage_pbar_rev(const Lane<NC, NG>& ln, double2* slot, int K) {
  constexpr int NE = NC * NC;
  if (!ln.lds_ok) return;
#pragma unroll
  for (int n = 0; n < NG; n++) slot[n * NE + ln.e] = (NG - 1 - n == K && ln.r == ln.c) ? one2() : zero2();
}

// Full Cauchy product of slot A and slot B into registers: C_n, n = 0..NG-1.
template <int NC, int NG>
__device__ __forceinline__ void conv(const Lane<NC, NG>& ln, double2 (&C)[NG]) {
#pragma unroll
  for (int k = 0; k < NG; k++) C[k] = zero2();
  __syncthreads();
  mul<NC, NG, 0>(ln, zero2(), zero2(), false, C, NG - 1, false, nullptr, false, 1.0);
  __syncthreads();
}

template <int NC, int NG>
__global__ __launch_bounds__(64, 1) void k_plaq_adjoint(Geo g, const double2* U, double2* G6, int K) {
  using LN = Lane<NC, NG>;
  __shared__ double2 lds[LN::LDS + NC * NC];
  // a third slot per link holds T1 between its product and its last use,
  // where it still leaves room for four workgroups per CU (up to 13 orders
  // of SU(3)); above that T1 stays in registers
  constexpr bool T1_LDS = (LN::LDS + NC * NC + LN::LPB * LN::SLOT) * 16 * 4 <= 163840;
  __shared__ double2 lds_t1[T1_LDS ? LN::LPB * LN::SLOT : 1];
  LN ln(lds);
  double2* const slotC = lds_t1 + (T1_LDS ? ln.j * LN::SLOT : 0);
  const int b = blockIdx.x;
  const int pl = b % 6, sb = b / 6;
  // planes (0,1) (0,2) (0,3) (1,2) (1,3) (2,3)
  const int mu = pl < 3 ? 0 : (pl < 5 ? 1 : 2);
  const int nu = pl < 3 ? pl + 1 : (pl < 5 ? pl - 1 : 3);
  const int x = min(sb * LN::LPB + ln.j, g.V - 1);
  const bool ok = ln.lds_ok && (sb * LN::LPB + ln.j) < g.V;
  const long V = g.V, P = 4 * V;
  const int xpmu = shift_site(g, x, mu, 1), xpnu = shift_site(g, x, nu, 1);
  const long l1 = mu * V + x, l2 = nu * V + xpmu, l3 = mu * V + xpnu, l4 = nu * V + x;
  const int ng = g.nrt;  // orders of the field (<= NG)

  // The eight products of a plaquette, one per iteration (the product body
  // is emitted once), ordered so that only two series live in registers
  // across a product -- the current cotangent Tb and the product R (and T1
  // where it does not wait in LDS); links are staged straight from memory
  // and Pbar = delta_{kK} 1 is generated (inlined eight times with the links, T1, T2 and Pbar held in
  // registers, the kernel spilled 238 VGPRs at 13 orders):
  //   0: R = U1 U2                                 T1 = R
  //   1: R = T1 U3^dag                             (T2)
  //   2: T2^dag rev(Pbar)  -> U4bar = rev(.)^dag
  //   3: R = rev(Pbar) U4                          Tb = T2bar = rev(R)
  //   4: T1^dag rev(T2bar) -> U3bar = rev(.)^dag
  //   5: R = rev(T2bar) U3                         Tb = T1bar = rev(R)
  //   6: rev(T1bar) U2^dag -> U1bar = rev(.)
  //   7: U1^dag rev(T1bar) -> U2bar = rev(.)
  constexpr int NE = NC * NC;
  const int etr = ln.c * NC + ln.r;
  double2 Tb[NG], R[NG], T1[T1_LDS ? 1 : NG];
#pragma unroll
  for (int k = 0; k < NG; k++) Tb[k] = zero2();
  for (int st = 0; st < 8; st++) {
    asm volatile("" ::: "memory");
    __syncthreads();
    if (st == 0) {
      stage_ser(ln, ln.slotA, U, P, l1, ng, false);
      stage_ser(ln, ln.slotB, U, P, l2, ng, false);
    } else if (st == 1) {
      if constexpr (T1_LDS) {
        if (ln.lds_ok) {
#pragma unroll
          for (int k = 0; k < NG; k++) ln.slotA[k * NE + ln.e] = slotC[k * NE + ln.e];
        }
      } else {
        ln.stage(ln.slotA, T1);
      }
      stage_ser(ln, ln.slotB, U, P, l3, ng, true);
    } else if (st == 2) {
      ln.stage(ln.slotA, R, true);
      stage_pbar_rev(ln, ln.slotB, K);
    } else if (st == 3) {
      stage_pbar_rev(ln, ln.slotA, K);
      stage_ser(ln, ln.slotB, U, P, l4, ng, false);
    } else if (st == 4) {
      if constexpr (T1_LDS) {
        if (ln.lds_ok) {
#pragma unroll
          for (int k = 0; k < NG; k++) ln.slotA[k * NE + etr] = c_conj(slotC[k * NE + ln.e]);
        }
      } else {
        ln.stage(ln.slotA, T1, true);
      }
      ln.stage(ln.slotB, Tb, false, true);
    } else if (st == 5) {
      ln.stage(ln.slotA, Tb, false, true);
      stage_ser(ln, ln.slotB, U, P, l3, ng, false);
    } else if (st == 6) {
      ln.stage(ln.slotA, Tb, false, true);
      stage_ser(ln, ln.slotB, U, P, l2, ng, true);
    } else {
      stage_ser(ln, ln.slotA, U, P, l1, ng, true);
      ln.stage(ln.slotB, Tb, false, true);
    }
    conv(ln, R);
    if (st == 0) {
      if constexpr (T1_LDS) {
        if (ln.lds_ok) {
#pragma unroll
          for (int k = 0; k < NG; k++) slotC[k * NE + ln.e] = R[k];
        }
      } else {
#pragma unroll
        for (int k = 0; k < NG; k++) T1[k] = R[k];
      }
    } else if (st == 3 || st == 5) {
#pragma unroll
      for (int k = 0; k < NG; k++) Tb[k] = R[NG - 1 - k];
    } else if (st != 1) {
      // gradient of one link, rev(R) or its dagger: slot 2 * (index of the
      // other direction) + (0 for the link at the plaquette's corner, 1 for
      // the far one); formed in slot A, then stored
      const long slot = st == 2   ? (long)(2 * other_index(nu, mu) + 0) * P + l4
                        : st == 4 ? (long)(2 * other_index(mu, nu) + 1) * P + l3
                        : st == 6 ? (long)(2 * other_index(mu, nu) + 0) * P + l1
                                  : (long)(2 * other_index(nu, mu) + 1) * P + l2;
      ln.stage(ln.slotA, R, st == 2 || st == 4, true);
      __syncthreads();
      if (ok) {
#pragma unroll
        for (int k = 0; k < NG; k++)
          if (k < ng) G6[((long)k * NE + ln.e) * 6 * P + slot] = ln.slotA[k * NE + ln.e];
      }
    }
  }
}

// Per link: G = sum of the six slots; Sigma_m = G_{K-m}^dag (K = ng-1, the
// field's top order); D = tau reH(U Sigma) per order 0..ng-1.  Gout / Dout
// may be null.
template <int NC, int NG>
__global__ __launch_bounds__(64, 1) void k_adjoint_drift(Geo g, const double2* U, const double2* G6,
                                                         double2* Gout, double2* Dout, double tau) {
  constexpr int NE = NC * NC;
  using LN = Lane<NC, NG>;
  __shared__ double2 lds[LN::LDS + NE];
  LN ln(lds);
  const int b = blockIdx.x;
  const int mu = b & 3, sb = b >> 2;
  const int x = min(sb * LN::LPB + ln.j, g.V - 1);
  const bool ok = ln.lds_ok && (sb * LN::LPB + ln.j) < g.V;
  const long V = g.V, P = 4 * V, idx = mu * V + x;
  const int ng = g.nrt;
  double2 Gs[NG], X[NG];
  // the link's own series first, then the six slots order by order: every
  // load is independent of the sums, so they are all in flight together
  // (the slots are summed in ascending order, as the gradient is defined)
  if (Dout) load_ser(ln, U, P, idx, ng, X);
#pragma unroll
  for (int k = 0; k < NG; k++) {
    double2 sum = zero2();
    if (k < ng) {
      const double2* gk = G6 + ((long)k * NE + ln.e) * 6 * P + idx;
#pragma unroll
      for (int slot = 0; slot < 6; slot++) sum = c_add(sum, gk[(long)slot * P]);
    }
    Gs[k] = sum;
  }
  if (Gout && ok)
#pragma unroll
    for (int k = 0; k < NG; k++)
      if (k < ng) Gout[((long)k * NE + ln.e) * P + idx] = Gs[k];
  if (!Dout) return;
  ln.stage(ln.slotA, X);
  // Sigma into slot B: order m is G_{ng-1-m}^dag (this lane writes the
  // conjugate of its element at the transposed position), zero above ng-1
  if (ln.lds_ok) {
    const int pos = ln.c * NC + ln.r;
#pragma unroll
    for (int m = 0; m < NG; m++) ln.slotB[m * NE + pos] = zero2();
#pragma unroll
    for (int k = 0; k < NG; k++)
      if (k < ng) ln.slotB[(ng - 1 - k) * NE + pos] = c_conj(Gs[k]);
  }
  double2 W[NG];
  conv(ln, W);
  ln.stage(ln.slotA, W);
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NG; k++) X[k] = c_scale(tau, reh_elem(ln, ln.slotA, k));
  if (ok)
#pragma unroll
    for (int k = 0; k < NG; k++)
      if (k < ng) Dout[((long)k * NE + ln.e) * P + idx] = X[k];
}

// host image (4V x NG*NE complex, link order 4*site+mu) <-> SoA
__global__ void k_g_aos_to_soa(const double2* aos, int S, int V, double2* soa) {
  const long P = 4L * V, n = P * S;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    const long j = t / P, idx = t % P;  // j = k*NE + e
    const long site = idx % V, mu = idx / V;
    soa[t] = aos[(4 * site + mu) * S + j];
  }
}
__global__ void k_g_soa_to_aos(const double2* soa, int S, int V, double2* aos) {
  const long P = 4L * V, n = P * S;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    const long link = t / S, j = t % S;
    const long site = link / 4, mu = link % 4;
    aos[t] = soa[j * P + mu * V + site];
  }
}

struct AdjointSet {
  int nc, ng;
  int lpb;  // links (plaquettes) per workgroup of these kernels: Lane<NC, NG>::LPB
  void (*plaq)(int nblk, Geo g, const double2* U, double2* G6, int K);
  void (*drift)(int nblk, Geo g, const double2* U, const double2* G6, double2* Gout, double2* Dout, double tau);
};

template <int NC, int NG>
AdjointSet adjoint_set() {
  AdjointSet s;
  s.nc = NC;
  s.ng = NG;
  s.lpb = Lane<NC, NG>::LPB;
  s.plaq = [](int nblk, Geo g, const double2* U, double2* G6, int K) {
    hipLaunchKernelGGL((k_plaq_adjoint<NC, NG>), dim3(nblk), dim3(64), 0, 0, g, U, G6, K);
  };
  s.drift = [](int nblk, Geo g, const double2* U, const double2* G6, double2* Gout, double2* Dout, double tau) {
    hipLaunchKernelGGL((k_adjoint_drift<NC, NG>), dim3(nblk), dim3(64), 0, 0, g, U, G6, Gout, Dout, tau);
  };
  return s;
}

// compiled orders (NG = ptord + 1, ascending); a field of ng orders runs on
// the smallest instance with NG >= ng
AdjointSet find_adjoint(int nc, int ng) {
  const AdjointSet all[] = {adjoint_set<3, 5>(),  adjoint_set<3, 7>(),  adjoint_set<3, 9>(),  adjoint_set<3, 13>(),
                            adjoint_set<3, 17>(), adjoint_set<3, 21>(), adjoint_set<2, 5>(),  adjoint_set<2, 7>(),
                            adjoint_set<2, 9>(),  adjoint_set<2, 13>(), adjoint_set<2, 17>(), adjoint_set<2, 21>()};
  for (const auto& s : all)
    if (s.nc == nc && s.ng >= ng) return s;
  return AdjointSet{0, 0, 0, nullptr, nullptr};
}

int grid_n(long n) {
  long gsz = (n + 255) / 256;
  return (int)(gsz > 65536 ? 65536 : (gsz < 1 ? 1 : gsz));
}

}  // namespace
}  // namespace nspt

// Scratch of nspt_adjoint_drift_dev: the six gradient slots of every link
// (6 x 4V series of ptord+1 matrices), in a static device array, which each
// GPU the library runs on holds until the process ends.  Host code reaches
// it only through the address hipGetSymbolAddress returns on that device.
constexpr size_t ADJ_SCRATCH_BYTES = size_t(192) << 20;
static __device__ __align__(256) double2 scratch[ADJ_SCRATCH_BYTES / sizeof(double2)];

namespace nspt {
namespace {

// Geometry of an adjoint call: 0, or -1 (bad extents or order) / -2 (order
// above the largest compiled instance).
int adjoint_geometry(const nspt_geom* geom, AdjointSet* set, Geo* g) {
  const int NE = geom->nc * geom->nc, NG = geom->ptord + 1;
  if (geom->ptord < 1) return -1;
  *set = find_adjoint(geom->nc, NG);
  if (!set->plaq) return -2;
  long V = 1;
  for (int d = 0; d < 4; d++) {
    if (geom->sz[d] < 2 || geom->sz[d] > 32) return -1;
    V *= geom->sz[d];
  }
  *g = Geo{};
  for (int d = 0; d < 4; d++) g->L[d] = geom->sz[d];
  g->st[3] = 1;
  g->st[2] = g->L[3];
  g->st[1] = g->L[2] * g->L[3];
  g->st[0] = g->L[1] * g->L[2] * g->L[3];
  g->V = (int)V;
  g->nrt = NG;
  (void)NE;
  return 0;
}

// Device address of `scratch` on the current device, looked up once per
// device (the array's host-side name is no device address).
constexpr int MAX_DEVICES = 64;
double2* g_scratch[MAX_DEVICES] = {};
std::mutex g_scratch_mu;

int scratch_on_current_device(double2** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) return -3;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return -3;
    g_scratch[dev] = static_cast<double2*>(p);
  }
  *out = g_scratch[dev];
  return 0;
}

}  // namespace
}  // namespace nspt

extern "C" int nspt_adjoint_drift_dev(const nspt_geom* geom, const double* U, double eps, double* grad,
                                      double* drift) {
  using namespace nspt;
  if (!geom) return -1;
  AdjointSet set;
  Geo g;
  const int rcg = adjoint_geometry(geom, &set, &g);
  if (rcg) return rcg;
  const int nc = geom->nc, NE = nc * nc, N = geom->ptord;
  const long V = g.V, P = 4 * V, S = (long)(N + 1) * NE;
  if ((size_t)(6 * S * P) * sizeof(double2) > ADJ_SCRATCH_BYTES) return -4;
  if (!U) return -1;
  double2* g6 = nullptr;
  if (scratch_on_current_device(&g6)) return -3;
  const long nb = (V + set.lpb - 1) / set.lpb;
  // every slot of every link is written by its plaquette: nothing to clear
  set.plaq((int)(6 * nb), g, reinterpret_cast<const double2*>(U), g6, N);
  set.drift((int)(4 * nb), g, reinterpret_cast<const double2*>(U), g6, reinterpret_cast<double2*>(grad),
            reinterpret_cast<double2*>(drift), -eps / (2.0 * nc));
  if (hipGetLastError() != hipSuccess) return -3;
  return hipStreamSynchronize(nullptr) == hipSuccess ? 0 : -3;
}

extern "C" int nspt_adjoint_drift(const nspt_geom* geom, const double* U, double eps, double* grad, double* drift) {
  using namespace nspt;
  if (!geom || !U) return -1;
  AdjointSet set;
  Geo g;
  const int rcg = adjoint_geometry(geom, &set, &g);
  if (rcg) return rcg;  // any order up to 20 (instances 4, 6, 8, 12, 16, 20)
  const int nc = geom->nc, N = geom->ptord, NG = N + 1, NE = nc * nc;
  const long V = g.V;
  const long nb = (V + set.lpb - 1) / set.lpb, P = 4 * V, S = (long)NG * NE;
  double2 *u = nullptr, *aos = nullptr, *g6 = nullptr, *go = nullptr, *dout = nullptr;
  int rc = 0;
  auto ok = [&](hipError_t e) {
    if (e != hipSuccess && rc == 0) rc = -3;
    return rc == 0;
  };
  if (ok(hipMalloc(&u, sizeof(double2) * S * P)) && ok(hipMalloc(&aos, sizeof(double2) * S * P)) &&
      ok(hipMalloc(&g6, sizeof(double2) * 6 * S * P)) && ok(hipMalloc(&go, sizeof(double2) * S * P)) &&
      ok(hipMalloc(&dout, sizeof(double2) * S * P)) &&
      ok(hipMemcpy(aos, U, sizeof(double2) * S * P, hipMemcpyHostToDevice))) {
    hipLaunchKernelGGL(k_g_aos_to_soa, dim3(grid_n(S * P)), dim3(256), 0, 0, (const double2*)aos, (int)S, (int)V, u);
    set.plaq((int)(6 * nb), g, u, g6, N);
    set.drift((int)(4 * nb), g, u, g6, go, dout, -eps / (2.0 * nc));
    ok(hipGetLastError());
    if (rc == 0 && grad) {
      hipLaunchKernelGGL(k_g_soa_to_aos, dim3(grid_n(S * P)), dim3(256), 0, 0, (const double2*)go, (int)S, (int)V, aos);
      ok(hipMemcpy(grad, aos, sizeof(double2) * S * P, hipMemcpyDeviceToHost));
    }
    if (rc == 0 && drift) {
      hipLaunchKernelGGL(k_g_soa_to_aos, dim3(grid_n(S * P)), dim3(256), 0, 0, (const double2*)dout, (int)S, (int)V,
                         aos);
      ok(hipMemcpy(drift, aos, sizeof(double2) * S * P, hipMemcpyDeviceToHost));
    }
    ok(hipDeviceSynchronize());
  }
  for (double2* p : {u, aos, g6, go, dout})
    if (p) (void)hipFree(p);
  return rc;
}
