// nspt_table.hpp -- arguments of one replica for the replica-table launches
// of the step kernels: one launch per stage covers every replica of a batch,
// blockIdx.y selecting the replica, whose arguments (fields, geometry,
// scalars, block count nblk) the kernel reads from a device table.
#pragma once
#include "nspt_device.hpp"
#include "nspt_internal.hpp"

namespace nspt {

struct DriftArgs {
  Geo g;
  SerField U;
  SerField D;
  SerField D1;
  double tau_g;
  double rk2c;
  int mode;
  double* obs_part;
  SerField R;
  double c1;
  int nblk;
};
struct PlaqArgs {
  Geo g;
  SerField U;
  SerField P;
  int nblk;
};
struct DriftPArgs {
  Geo g;
  SerField U;
  SerField P;
  SerField D;
  SerField D1;
  double tau_g;
  double rk2c;
  int mode;
  double* obs_part;
  SerField R;
  double c1;
  int nblk;
};
struct UpdateArgs {
  Geo g;
  SerField D;
  const double2* eta;
  double stau;
  SerField Ubase;
  SerField Uout;
  SerField L;
  double2* mm_part;
  int nblk;
};
struct ZeroModeArgs {
  Geo g;
  SerField L;
  const double2* MM;
  SerField U;
  SerField X;
  int nblk;
};
struct GexpArgs {
  Geo g;
  const double2* fb;
  SerField G;
  int nblk;
};
struct TransformArgs {
  Geo g;
  SerField G;
  SerField U;
  SerField Xexp;
  const double2* MM;
  int nblk;
};

}  // namespace nspt
