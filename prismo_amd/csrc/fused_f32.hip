// FP32 instantiations of the fused Yee step (see fused_kernel.h).
#include "fused_kernel.h"
#include "fused2_kernel.h"

namespace pm {

int launch_yee_fused_f32(const KGrid& g, int ck, const FusedBufs& fb, int vec, int w, int ty,
                         int kchunk, int order, int xs, hipStream_t s) {
  KFz z;
  if (!make_kfz(g, fb, (int)sizeof(float), &z)) return -1;
  if (ck == PM_COEFF_FULL)
    return launch_fused_ck<float, PM_COEFF_FULL>(g, z, vec, w, ty, kchunk, order, xs, s);
  if (ck == PM_COEFF_CB)
    return launch_fused_ck<float, PM_COEFF_CB>(g, z, vec, w, ty, kchunk, order, xs, s);
  return launch_fused_ck<float, PM_COEFF_CB_F16>(g, z, vec, w, ty, kchunk, order, xs, s);
}

bool fused2_grid_ok(const KGrid& g) { return g.pml[0] + g.pml[1] <= kF2XSlab; }

bool fused2_supported(int vec, int w, int ty) {
  return ty == kF2TyIn && ((vec == 2 && w == 2) || (vec == 1 && w == 2));
}

int launch_yee_fused2_f32(const KGrid& g, int ck, const FusedBufs& fb, int vec, int w, int ty,
                          int kchunk, int order, int64_t src_row, int n_src, const KSrc* src,
                          hipStream_t s, hipStream_t s2, hipEvent_t fork, hipEvent_t join) {
  KFz2 z2{};
  if (!fused2_grid_ok(g)) return -1;
  if (!make_kfz(g, fb, (int)sizeof(float), &z2.z)) return -1;
  z2.src_row = src_row;
  z2.n_src = n_src;
  z2.src = src;
  // put every x CPML cell of a row in its first wave when the second wave's
  // half-row fits between the layers (an even split point keeps each lane's
  // two cells aligned)
  {
    const int rw = 64 * vec * w, half = rw / 2;
    int xa = std::min(half, g.nx - g.pml[1] - half);
    xa -= xa % (2 * vec);
    z2.xa = (w == 2 && g.has_pml && xa >= g.pml[0] && xa > 0 && xa + half <= g.nx - g.pml[1])
                ? xa : 0;
  }
  if (ty != kF2TyIn) return -1;
  const F2Streams st{s, s2, fork, join};
  if (ck == PM_COEFF_CB) {
    if (vec == 2 && w == 2) return launch_fused2_one<PM_COEFF_CB, 2, 2>(g, z2, kchunk, order, st);
    if (vec == 1 && w == 2) return launch_fused2_one<PM_COEFF_CB, 1, 2>(g, z2, kchunk, order, st);
  } else if (ck == PM_COEFF_CB_F16) {
    if (vec == 2 && w == 2)
      return launch_fused2_one<PM_COEFF_CB_F16, 2, 2>(g, z2, kchunk, order, st);
    if (vec == 1 && w == 2)
      return launch_fused2_one<PM_COEFF_CB_F16, 1, 2>(g, z2, kchunk, order, st);
  }
  return -1;
}

bool fused_addressable(const KGrid& g, const FusedBufs& fb, int esz) {
  KFz z;
  return make_kfz(g, fb, esz, &z);
}

}  // namespace pm
