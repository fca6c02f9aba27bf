// nspt_internal.hpp -- types shared by the kernel instances and the host
// side of the library (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>

namespace nspt {

struct SerField;

// Lattice geometry as the kernels see it.  st[d] is the site stride of
// direction d (x3 fastest, lattice.h:37).
struct Geo {
  int L[4];
  int st[4];
  int V;
  int nrt;    // runtime order (PTORD)
  int nb_mu;  // link blocks per direction = ceil(V / links-per-block)
};

// Launchers of one (NC, N) instance.  nblk is the grid size in waves.
struct KernelSet {
  int nc = 0, nmax = 0, lpb = 0;
  void (*drift)(int nblk, hipStream_t, Geo, SerField U, SerField D, SerField D1, double tau_g, double rk2c, int mode,
                double* obs) = nullptr;
  void (*update)(int nblk, hipStream_t, Geo, SerField D, const double2* eta, double stau, SerField Ubase,
                 SerField Uout, SerField L, double2* mm) = nullptr;
  void (*logf)(int nblk, hipStream_t, Geo, SerField U, SerField L) = nullptr;
  void (*zero_mode)(int nblk, hipStream_t, Geo, SerField L, const double2* MM, SerField U, SerField X) = nullptr;
  void (*gexp)(int nblk, hipStream_t, Geo, const double2* fb, SerField G) = nullptr;
  void (*transform)(int nblk, hipStream_t, Geo, SerField G, SerField U) = nullptr;
};

// One function per compiled instance (nspt_inst.hip built once per (NC, N)).
#ifdef NSPT_SINGLE_INSTANCE
#define NSPT_INSTANCES(X) X(3, 20)
#else
#define NSPT_INSTANCES(X) \
  X(3, 4) X(3, 6) X(3, 8) X(3, 12) X(3, 16) X(3, 20) X(3, 24) \
  X(2, 4) X(2, 6) X(2, 8) X(2, 12) X(2, 16) X(2, 20) X(2, 24)
#endif

#define NSPT_DECL_SET(nc, n) KernelSet kernel_set_##nc##_##n();
NSPT_INSTANCES(NSPT_DECL_SET)
#undef NSPT_DECL_SET

}  // namespace nspt
