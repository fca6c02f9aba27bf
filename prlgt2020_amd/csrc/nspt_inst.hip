// nspt_inst.hip -- one (NC, N, tile edge, links per wave, waves per SIMD)
// instance of the link/site kernels.  Compiled once per instance with
// -DINST_NC=<nc> -DINST_N=<n> -DINST_QT=<edge> -DINST_LPW=<links per wave,
// 0 for the default> -DINST_WPE=<waves per SIMD>, and again with
// -DINST_BIG=1 for fields past 2 GiB.
#ifndef INST_QT
#define INST_QT 4
#endif
#ifndef INST_LPW
#define INST_LPW 0
#endif
#ifndef INST_WPE
#define INST_WPE 1
#endif
#ifndef NSPT_QT
#define NSPT_QT INST_QT
#endif
#define NSPT_LPW INST_LPW
#define NSPT_WAVES_PER_EU INST_WPE
#ifndef INST_BIG
#define INST_BIG 0
#endif
#define NSPT_BIG_FIELDS INST_BIG
#include "nspt_kernels.hpp"

#if INST_BIG
#define NSPT_CAT2(a, b, q, l, w) kernel_set_##a##_##b##_q##q##_l##l##_w##w##_big
#else
#define NSPT_CAT2(a, b, q, l, w) kernel_set_##a##_##b##_q##q##_l##l##_w##w
#endif
#define NSPT_CAT(a, b, q, l, w) NSPT_CAT2(a, b, q, l, w)

namespace nspt {
KernelSet NSPT_CAT(INST_NC, INST_N, INST_QT, INST_LPW, INST_WPE)() { return make_kernel_set<INST_NC, INST_N>(); }
}  // namespace nspt
