// FP64 instantiations of the fused Yee step (see fused_kernel.h).
#include "fused_kernel.h"

namespace pm {

int launch_yee_fused_f64(const KGrid& g, int ck, const FusedBufs& fb, int vec, int w, int ty,
                         int kchunk, int order, int xs, hipStream_t s) {
  KFz z;
  if (!make_kfz(g, fb, (int)sizeof(double), &z)) return -1;
  if (ck == PM_COEFF_FULL)
    return launch_fused_ck<double, PM_COEFF_FULL>(g, z, vec, w, ty, kchunk, order, xs, s);
  return launch_fused_ck<double, PM_COEFF_CB>(g, z, vec, w, ty, kchunk, order, xs, s);
}

bool fused_supported(int precision, int ck, int vec, int w, int ty, int xs) {
  const size_t t = precision == PM_FP64 ? 8 : 4;
  if (ty < 2 || ty > 7 || ty == 6) return false;
  if (ty > 4 && !(vec == 1 && w == 2)) return false;
  if (xs == 2) {
    if (!fused_split_built((int)t, ck, vec, w)) return false;
  } else {
    const bool pair = (vec == 1 && (w == 2 || w == 4)) ||
                      (vec == 2 && (w == 1 || w == 2 || w == 4)) || (vec == 4 && (w == 1 || w == 2));
    if (xs != 1 || !pair) return false;
  }
  return 64 * w * (ty + 1) <= 1024 && (size_t)9 * (ty + 1) * 64 * vec * w * t <= 65536;
}

}  // namespace pm
