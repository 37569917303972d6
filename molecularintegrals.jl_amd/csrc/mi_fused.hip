// mi_fused.hip -- fused register kernel hgp_rf<g>_kernel of fusion group g
// (compiled once per group with -DMI_FUSED_GROUP=g): one launch over the
// register-kernel class launches of that group in a plan (each a segment of
// the grid).  Each group is a translation unit of its own because the kernel
// inlines every body of the group; mi_kernels.hip holds the single-class
// kernels hgp_r_<class> built from the same bodies.
#ifndef MI_FUSED_GROUP
#error "build with -DMI_FUSED_GROUP=<group>"
#endif
#include "mi_device.h"
#include "mi_kernels.h"

#define MI_R_BODIES_ONLY
#include "gen/hgp_rkernels.inc"
#include "gen/hgp_rfused.inc"

#define MI_CAT2(a, b) a##b
#define MI_CAT(a, b) MI_CAT2(a, b)

namespace mi {

// launch_fused_register_g<MI_FUSED_GROUP>
hipError_t MI_CAT(launch_fused_register_g, MI_FUSED_GROUP)(const EriArgs& args, const FusedSegs& F, int nblocks,
                                                          int wpb, hipStream_t stream) {
  void* kargs[] = {const_cast<EriArgs*>(&args), const_cast<FusedSegs*>(&F)};
  return hipLaunchKernel(kFusedRegisterKernel, dim3(nblocks), dim3(64 * wpb), kargs, 0, stream);
}

}  // namespace mi
