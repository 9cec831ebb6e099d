// Instantiations of one stage kernel (k_vel or k_str, FW_KERNEL 0 / 1) at
// one tile shape (FW_TILE 0..3) over the 8 F/B patterns x 3 RK kinds.  The
// build compiles this file once per (kernel, shape) into separate objects so
// that the 192 stage kernels compile in parallel; each object exports
// fw::stage_fn_{v,s}<shape>(pattern, kind), the host stub of the kernel.
#include "fw_stage.h"

#if !defined(FW_KERNEL) || !defined(FW_TILE)
#error "compile with -DFW_KERNEL=0|1 -DFW_TILE=0..3"
#endif

#define FW_CAT2(a, b) a##b
#define FW_CAT(a, b) FW_CAT2(a, b)
#if FW_KERNEL == 0
#define FW_FN FW_CAT(stage_fn_v, FW_TILE)
#else
#define FW_FN FW_CAT(stage_fn_s, FW_TILE)
#endif

namespace fw {

template <int P, int K>
static StageFn one() {
  constexpr int XF = (P >> 2) & 1, YF = (P >> 1) & 1, ZF = P & 1, B = TILE_BYT[FW_TILE], R = TILE_RY[FW_TILE];
#if FW_KERNEL == 0
  return k_vel<XF, YF, ZF, K, B, R>;
#else
  return k_str<XF, YF, ZF, K, B, R>;
#endif
}
#define FW_ROW(P) {one<P, 0>(), one<P, 1>(), one<P, 2>()}

StageFn FW_FN(int pat, int kind) {
  static const StageFn T[8][3] = {FW_ROW(0), FW_ROW(1), FW_ROW(2), FW_ROW(3),
                                  FW_ROW(4), FW_ROW(5), FW_ROW(6), FW_ROW(7)};
  return T[pat & 7][kind];
}

}  // namespace fw
