// nspt_inst.hip -- one (NC, N) instance of the link/site kernels.  Compiled
// once per instance with -DINST_NC=<nc> -DINST_N=<n>.
#include "nspt_kernels.hpp"

#define NSPT_CAT2(a, b, c) kernel_set_##a##_##b##c
#define NSPT_CAT(a, b) NSPT_CAT2(a, b, )

namespace nspt {
KernelSet NSPT_CAT(INST_NC, INST_N)() { return make_kernel_set<INST_NC, INST_N>(); }
}  // namespace nspt
