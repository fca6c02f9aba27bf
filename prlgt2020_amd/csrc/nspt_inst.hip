// nspt_inst.hip -- one (NC, N, tile edge) instance of the link/site
// kernels.  Compiled once per instance with -DINST_NC=<nc> -DINST_N=<n>
// -DINST_QT=<edge>, and again with -DINST_BIG=1 for fields past 2 GiB.
#ifndef INST_QT
#define INST_QT 4
#endif
#ifndef NSPT_QT
#define NSPT_QT INST_QT
#endif
#ifndef INST_BIG
#define INST_BIG 0
#endif
#define NSPT_BIG_FIELDS INST_BIG
#include "nspt_kernels.hpp"

#if INST_BIG
#define NSPT_CAT2(a, b, q) kernel_set_##a##_##b##_q##q##_big
#else
#define NSPT_CAT2(a, b, q) kernel_set_##a##_##b##_q##q
#endif
#define NSPT_CAT(a, b, q) NSPT_CAT2(a, b, q)

namespace nspt {
KernelSet NSPT_CAT(INST_NC, INST_N, INST_QT)() { return make_kernel_set<INST_NC, INST_N>(); }
}  // namespace nspt
