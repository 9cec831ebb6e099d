// Instantiations of the stage kernel at one tile shape (FW_TILE 0..3) for
// half of the 8 F/B patterns (FW_PHALF 0: patterns 0-3, 1: 4-7) x the 4 RK
// forms, on the spherical grid (FW_CURV 0) or the terrain-following grid of
// mode C (FW_CURV 1).  The build compiles this file once per (shape, half,
// grid) into separate objects so that the 256 stage kernels compile in
// parallel; each object exports fw::stage_fn_<shape>_<half>(pattern, form)
// (stage_fnc_ for mode C), the host stub of the kernel.
#include "fw_stage.h"

#if !defined(FW_TILE) || !defined(FW_PHALF) || !defined(FW_CURV)
#error "compile with -DFW_TILE=0..3 -DFW_PHALF=0|1 -DFW_CURV=0|1"
#endif

#define FW_CAT4(a, b, c, d) a##b##c##d
#if FW_CURV
#define FW_NAME(t, h) FW_CAT4(stage_fnc_, t, _, h)
#else
#define FW_NAME(t, h) FW_CAT4(stage_fn_, t, _, h)
#endif

namespace fw {

StageFn FW_NAME(FW_TILE, FW_PHALF)(int pat, int form);

template <int P, int F>
static StageFn one() {
  constexpr int XF = (P >> 2) & 1, YF = (P >> 1) & 1, ZF = P & 1, B = TILE_BYT[FW_TILE], R = TILE_RY[FW_TILE];
  return k_stage<XF, YF, ZF, F, B, R, FW_CURV != 0>;
}
#define FW_ROW(P) {one<P, 0>(), one<P, 1>(), one<P, 2>(), one<P, 3>()}

StageFn FW_NAME(FW_TILE, FW_PHALF)(int pat, int form) {
  constexpr int P0 = 4 * FW_PHALF;
  static const StageFn T[4][4] = {FW_ROW(P0), FW_ROW(P0 + 1), FW_ROW(P0 + 2), FW_ROW(P0 + 3)};
  return T[(pat - P0) & 3][form & 3];
}

}  // namespace fw
