// pxr_fused_mxfx.hip -- mixed-precision instantiations of the fused gather + push +
// deposit kernel, LDS current accumulated in 64-bit fixed point, with its
// cut-block deposition kernel (compiled apart from the other variants to
// shorten the build).
#include "pxr_fused.h"
#include "pxr_runtime.h"

namespace pxr {

static const int FUSED_BS = 256;

template <int JACC>
static int launch(int key, const FusedArgs &a, int nwg, size_t lds, hipStream_t st) {
  const dim3 grd(nwg), blk(FUSED_BS);
#define PXR_FUSED_CASE(N_, LO_, P_, D_)                                                          \
  case N_ * 100 + (LO_ ? 10 : 0) + P_ * 2 + (D_ ? 1 : 0):                                        \
    if (lds > 65536)                                                                             \
      PXR_HIP(hipFuncSetAttribute((const void *)&k_fused<FUSED_BS, N_, LO_, P_, D_, true, JACC>, \
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));       \
    hipLaunchKernelGGL((k_fused<FUSED_BS, N_, LO_, P_, D_, true, JACC>), grd, blk, lds, st, a);  \
    if (a.phase & PHASE_CUT) {                                                                   \
      if (lds > 65536)                                                                           \
        PXR_HIP(hipFuncSetAttribute((const void *)&k_fused_cut<FUSED_BS, N_, LO_, P_, D_>,       \
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));     \
      hipLaunchKernelGGL((k_fused_cut<FUSED_BS, N_, LO_, P_, D_>), grd, blk, lds, st, a);        \
    }                                                                                            \
    hipLaunchKernelGGL((k_fused_overflow<FUSED_BS, N_, LO_, P_, D_, true>), dim3(2 * rt().num_cus), \
                       blk, 0, st, a);                                                           \
    break;
  switch (key) {
    PXR_FUSED_CASE(1, true, 0, 0) PXR_FUSED_CASE(1, true, 0, 3) PXR_FUSED_CASE(1, true, 1, 0)
    PXR_FUSED_CASE(1, true, 1, 3) PXR_FUSED_CASE(1, false, 0, 0) PXR_FUSED_CASE(1, false, 0, 3)
    PXR_FUSED_CASE(1, false, 1, 0) PXR_FUSED_CASE(1, false, 1, 3)
    PXR_FUSED_CASE(2, true, 0, 0) PXR_FUSED_CASE(2, true, 0, 3) PXR_FUSED_CASE(2, true, 1, 0)
    PXR_FUSED_CASE(2, true, 1, 3) PXR_FUSED_CASE(2, false, 0, 0) PXR_FUSED_CASE(2, false, 0, 3)
    PXR_FUSED_CASE(2, false, 1, 0) PXR_FUSED_CASE(2, false, 1, 3)
    PXR_FUSED_CASE(3, true, 0, 0) PXR_FUSED_CASE(3, true, 0, 3) PXR_FUSED_CASE(3, true, 1, 0)
    PXR_FUSED_CASE(3, true, 1, 3) PXR_FUSED_CASE(3, false, 0, 0) PXR_FUSED_CASE(3, false, 0, 3)
    PXR_FUSED_CASE(3, false, 1, 0) PXR_FUSED_CASE(3, false, 1, 3)
    default:
      return set_error(PXRHIP_ERR_ARG, "unsupported kernel combination");
  }
#undef PXR_FUSED_CASE
  PXR_LAUNCHED();
  return PXRHIP_OK;
}

int launch_fused_mixed_fx(int key, const FusedArgs &a, int nwg, size_t lds, hipStream_t st) {
  return launch<3>(key, a, nwg, lds, st);
}

}  // namespace pxr
