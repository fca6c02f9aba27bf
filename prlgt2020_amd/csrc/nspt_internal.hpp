// nspt_internal.hpp -- types shared by the kernel instances and the host
// side of the library (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>

namespace nspt {

struct SerField;
struct DriftArgs;
struct PlaqArgs;
struct DriftPArgs;
struct UpdateArgs;
struct ZeroModeArgs;
struct GexpArgs;
struct TransformArgs;

// Lattice geometry as the kernels see it.  st[d] is the site stride of
// direction d (x3 fastest, lattice.h:37).
struct Geo {
  int L[4];
  int st[4];
  int V;
  int nrt;    // runtime order (PTORD)
  int nb_mu;  // link blocks per direction = ceil(V / links-per-block)
  // Link subset of a checkerboard pass (0: every link): sub = 1 + 4 parity
  // + mu selects the links (x, mu) with (x0+x1+x2+x3) mod 2 = parity, in
  // nb_half blocks (ceil(V/2 / links-per-block)); needs L_3 even.
  int sub;
  int nb_half;
  // Twisted boundary conditions: bit d of twist set means
  // U(x + L_d e_d) = W_d U(x) W_d^dag, W_d described by tw (device memory).
  int twist;
  const struct TwistTab* tw;
};

// Monomial twist matrices: row i of W_d holds its one nonzero entry
// ph[d][i] in column (perm[d] >> 2i) & 3; bits 8 + 2i of perm[d] hold the
// inverse permutation.
struct TwistTab {
  double2 ph[4][3];
  unsigned perm[4];
};

// Launchers of one (NC, N) instance.  nblk is the grid size in waves.
struct KernelSet {
  int nc = 0, nmax = 0, lpb = 0, qt = 0;  // qt: order-tile edge of the products
  int lpw = 0;                            // links per wave asked of the build (0: the default, lpb)
  int wpe = 1;                            // waves per SIMD the build's register budget allows
  bool big = false;                       // built for fields past 2 GiB (NSPT_BIG_FIELDS)
  void (*drift)(int nblk, hipStream_t, Geo, SerField U, SerField D, SerField D1, double tau_g, double rk2c, int mode,
                double* obs, SerField R, double c1) = nullptr;
  // the plaquette field of U (6 per site, plane-major), and the drift of
  // `drift` formed from it (Jacobi passes only: g.sub = 0)
  void (*plaq_field)(int nblk, hipStream_t, Geo, SerField U, SerField P) = nullptr;
  void (*drift_p)(int nblk, hipStream_t, Geo, SerField U, SerField P, SerField D, SerField D1, double tau_g,
                  double rk2c, int mode, double* obs, SerField R, double c1) = nullptr;
  // sum of the rectangle staples of every link from U and its 2-link field
  // L2, and (obs2 != null) their per-block plaquette traces
  void (*rect)(int nblk, hipStream_t, Geo, SerField U, SerField L2, SerField R, double* obs2) = nullptr;
  // one checkerboard pass of the sequential gauge fixing (g.sub = 1 + 4 parity)
  void (*sgf)(int nblk, hipStream_t, Geo, SerField U, double alpha) = nullptr;
  void (*update)(int nblk, hipStream_t, Geo, SerField D, const double2* eta, double stau, SerField Ubase,
                 SerField Uout, SerField L, double2* mm) = nullptr;
  void (*logf)(int nblk, hipStream_t, Geo, SerField U, SerField L) = nullptr;
  void (*zero_mode)(int nblk, hipStream_t, Geo, SerField L, const double2* MM, SerField U, SerField X) = nullptr;
  void (*gexp)(int nblk, hipStream_t, Geo, const double2* fb, SerField G) = nullptr;
  void (*transform)(int nblk, hipStream_t, Geo, SerField G, SerField U, SerField Xexp,
                    const double2* MM) = nullptr;
  // the same kernels launched once for the replicas of a batch, each
  // replica's arguments in a device table (grid: gx blocks, a multiple of 8
  // covering the largest replica, times nrep)
  void (*drift_tab)(int gx, int nrep, hipStream_t, const DriftArgs* tab) = nullptr;
  void (*plaq_field_tab)(int gx, int nrep, hipStream_t, const PlaqArgs* tab) = nullptr;
  void (*drift_p_tab)(int gx, int nrep, hipStream_t, const DriftPArgs* tab) = nullptr;
  void (*update_tab)(int gx, int nrep, hipStream_t, const UpdateArgs* tab) = nullptr;
  void (*zero_mode_tab)(int gx, int nrep, hipStream_t, const ZeroModeArgs* tab) = nullptr;
  void (*gexp_tab)(int gx, int nrep, hipStream_t, const GexpArgs* tab) = nullptr;
  void (*transform_tab)(int gx, int nrep, hipStream_t, const TransformArgs* tab) = nullptr;
  void (*link_pair)(int nblk, hipStream_t, Geo, SerField A, SerField out, int k) = nullptr;
  void (*loops)(int nblk, hipStream_t, Geo, SerField U, SerField L2, SerField L4, double* part) = nullptr;
};

// One function per compiled instance (nspt_inst.hip built once per
// X(NC, N, tile edge, links per wave, waves per SIMD); links per wave 0 is
// the instance's own choice, links_per_wave()).  The default instance of an
// order has tile edge 4, links per wave 0 and one wave per SIMD; the SU(3)
// orders of the production configurations also come with tile edge 3, and
// N = 16, 20 with occupancy variants (4 links per wave at one and at two
// waves per SIMD), for the batch autotuner and for A/B runs.
#ifdef NSPT_SINGLE_INSTANCE
#ifndef NSPT_SI_N
#define NSPT_SI_N 20
#endif
#define NSPT_SI_APPLY(X, a, b, c, l, w) X(a, b, c, l, w)
#define NSPT_INSTANCES(X) NSPT_SI_APPLY(X, 3, NSPT_SI_N, 4, 0, 1)
#else
#define NSPT_INSTANCES(X)                                                                                           \
  X(3, 4, 4, 0, 1) X(3, 6, 4, 0, 1) X(3, 8, 4, 0, 1) X(3, 12, 4, 0, 1) X(3, 16, 4, 0, 1) X(3, 16, 3, 0, 1)           \
  X(3, 16, 4, 4, 1) X(3, 16, 4, 4, 2) X(3, 20, 4, 0, 1) X(3, 20, 3, 0, 1) X(3, 20, 4, 4, 2) X(3, 24, 4, 0, 1)        \
  X(3, 32, 4, 0, 1) X(2, 4, 4, 0, 1) X(2, 6, 4, 0, 1) X(2, 8, 4, 0, 1) X(2, 12, 4, 0, 1) X(2, 16, 4, 0, 1)            \
  X(2, 20, 4, 0, 1) X(2, 24, 4, 0, 1)
#endif

// Each instance comes in two builds: for fields whose order planes lie
// within 2 GiB of the field's base, and (_big) for larger ones.
#define NSPT_DECL_SET(nc, n, q, l, w)                   \
  KernelSet kernel_set_##nc##_##n##_q##q##_l##l##_w##w(); \
  KernelSet kernel_set_##nc##_##n##_q##q##_l##l##_w##w##_big();
NSPT_INSTANCES(NSPT_DECL_SET)
#undef NSPT_DECL_SET

}  // namespace nspt
