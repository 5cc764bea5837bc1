// FP32 instantiations of the fused Yee step (see fused_kernel.h).
#include "fused_kernel.h"

namespace pm {

int launch_yee_fused_f32(const KGrid& g, int ck, const FusedBufs& fb, int vec, int w, int ty,
                         int kchunk, int order, hipStream_t s) {
  KFz z;
  if (!make_kfz(g, fb, (int)sizeof(float), &z)) return -1;
  if (ck == PM_COEFF_FULL)
    return launch_fused_ck<float, PM_COEFF_FULL>(g, z, vec, w, ty, kchunk, order, s);
  if (ck == PM_COEFF_CB)
    return launch_fused_ck<float, PM_COEFF_CB>(g, z, vec, w, ty, kchunk, order, s);
  return launch_fused_ck<float, PM_COEFF_CB_F16>(g, z, vec, w, ty, kchunk, order, s);
}

bool fused_addressable(const KGrid& g, const FusedBufs& fb, int esz) {
  KFz z;
  return make_kfz(g, fb, esz, &z);
}

}  // namespace pm
