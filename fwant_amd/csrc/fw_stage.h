// Stage kernels of libfwant: the operator sweep Lx?_Ly?_Lz?
// (code/srcF/mod_macdrp.F90:790-962) fused with the RK update
// (mod_macdrp.F90:291-393), split by output:
//   k_vel  hVx hVy hVz from the six stresses (+ 1/rho)          -> 3 fields
//   k_str  hTxx .. hTyz from the three velocities (+ lambda, mu) -> 6 fields
// The RHS of a velocity depends on stresses only and vice versa, so the
// split adds no traffic: every input field is read once per stage by one of
// the two kernels.
//
// Work decomposition (both kernels): a 256-thread block (4 waves) owns a
// TW x TY tile of (i, j) columns over a chunk of k planes and marches k.
// Tiles are TW = 64 columns wide (a wave row is one 256-byte row segment),
// except the narrow tiles at the x edge of the volume (the remainder of the
// x extent after whole 64-column tiles, when it is at most 32 columns),
// which are TW = 16 columns wide: a wave then covers 4 rows of 16 columns,
// and no lanes idle on a 16-column remainder.
// The current k plane of every field that needs an x or y derivative is
// staged in LDS with its one-sided halo (F: -1..+3, B: -3..+1), double
// buffered so one barrier per plane suffices; z derivatives come from a
// 5-slot register window per column.  While plane k is computed, the loads
// of plane k+1 (window lead, halo) and the RK operands of plane k are in
// flight.  In a 64-wide tile wave w owns rows j0 + w + 4q (q < RY): every
// global access of a wave is 64 consecutive floats of one row.  Offsets are
// 32-bit relative to per-field base pointers.
#pragma once
#include "fw_common.h"

namespace fw {

constexpr int TX = 64;   // lanes per wave, columns of a wide tile
constexpr int TXN = 16;  // columns of a narrow tile

// Streamed operands (RK registers m and t, the stage output, the media) are
// touched once per stage: non-temporal accesses keep them from evicting the
// stencil planes whose halos neighbouring tiles re-read from L2.
#ifndef FW_STREAM_NT
#define FW_STREAM_NT 1
#endif
__device__ __forceinline__ float lds_nt(const float* p) {
#if FW_STREAM_NT
  return __builtin_nontemporal_load(p);
#else
  return *p;
#endif
}
__device__ __forceinline__ void st_nt(float* p, float v) {
#if FW_STREAM_NT
  __builtin_nontemporal_store(v, p);
#else
  *p = v;
#endif
}

// One CFS-PML face slab (x1 x2 y1 y2 z1 z2): inclusive bounds, extents and
// the auxiliary variables, 4 register sets (A B C T) x 6 variables x cnt.
// Variable order: x: Vx Vy Vz Txx Txy Txz; y: Vx Vy Vz Tyy Txy Tyz;
// z: Vx Vy Vz Tzz Txz Tyz (the psi of the derivative of that field).
struct FaceArgs {
  float* d;
  long cnt;
  int lo0, lo1, lo2, hi0, hi1, hi2, ex0, ex1, on;
};

struct StageArgs {
  const float* __restrict__ win;
  float* __restrict__ wout;
  const float* __restrict__ m;
  float* __restrict__ t;
  const float* __restrict__ lam;
  const float* __restrict__ mu;
  const float* __restrict__ rrho;
  const float* __restrict__ xi_x;
  const float* __restrict__ eta_y;
  const float* __restrict__ zeta_z;
  const float* __restrict__ rz;
  const float* __restrict__ rxsin;
  const float* __restrict__ xcot;
  const float* __restrict__ msurf;       // pre-image Tzz, Txz, Tyz at k = nk2 (3 planes)
  const int* __restrict__ fluidcol;      // 1 where the (i, j) column has a cell with mu <= SEIS_ZERO
  long fs, sy, sz, ps;
  int nx, ny;
  int ni1, ni2, nj1, nj2, nk1, nk2;
  int freenode;
  const int4* __restrict__ items;  // work items (i0 | ncol << 16, j0 | nrow << 16, k0, k1): a tile over a chunk of planes
  int nitems, p1, p2;  // item parts: narrow [0, p1), wide PML [p1, p2), wide interior
  float a, b;
  // CFS-PML: per-axis profiles 1/B, A + D/B, D/B (x: nx, y: ny, z: nz entries)
  const float* __restrict__ pb1;
  const float* __restrict__ pad;
  const float* __restrict__ pdd;
  int n1, setin, setout, pml;
  FaceArgs face[6];
  // batched virtual sources: nm members from slot0 in one launch; a
  // member's register sets start bstride floats after the previous
  // member's, its surface copy 3*ps after, its PML auxiliaries 24*cnt after
  // (per face)
  long bstride;
  int slot0, nm;
  int xg;  // items per XCD run (half the wide tiles of a row)
};

// Blocks are dealt round-robin to the 8 XCDs.  The (work item, member)
// pairs of a launch are ordered item-major (the members of one item run side
// by side on one XCD and share its media reads in L2), and the items by
// chunk, then tile row, then tile column.  Runs of xg consecutive pairs (xg:
// half the wide tiles of a row, times the members) go to one XCD, the next
// run to the next XCD: neighbouring tiles of a row share their x halos in one
// L2, and every XCD gets an even share of every kind of item.  (Giving each XCD
// one contiguous eighth of the list instead keeps more halos in one L2 but
// measures 4-5% slower: the eighths differ in cost.)  The item list holds
// three parts, dispatched in this order: the narrow edge tiles [0, p1), the
// wide PML-shell tiles [p1, p2) (the heavier ones), the wide interior tiles
// [p2, nitems); each part is mapped on its own.
struct ItemPos {
  int item, member;
};
__device__ __forceinline__ ItemPos item_pos(const StageArgs& g) {
  const int nm = g.nm, b = blockIdx.x;
  const int e1 = g.p1 * nm, e2 = g.p2 * nm, e3 = g.nitems * nm;
  const int base = b < e1 ? 0 : (b < e2 ? e1 : e2);
  const int end = b < e1 ? e1 : (b < e2 ? e2 : e3);
  const int n = end - base, bb = b - base, G = g.xg * nm, R = 8 * G, full = (n / R) * R;
  const int L = bb < full ? ((bb / 8) / G) * R + (bb % 8) * G + (bb / 8) % G : bb;
  return ItemPos{base / nm + L / nm, L % nm};
}

__device__ __forceinline__ long face_index(const FaceArgs& f, int i, int j, int k) {
  // a face holds fewer than 2^31 cells (checked at set-up)
  return (long)((i - f.lo0) + f.ex0 * ((j - f.lo1) + f.ex1 * (k - f.lo2)));
}
__device__ __forceinline__ bool in_face(const FaceArgs& f, int i, int j, int k) {
  return f.on && i >= f.lo0 && i <= f.hi0 && j >= f.lo1 && j <= f.hi1 && k >= f.lo2 && k <= f.hi2;
}
// slot: the member's slot in the batch
__device__ __forceinline__ float* aux_base(const FaceArgs& f, int slot) {
  return f.d + (long)slot * 24 * f.cnt;
}

// Auxiliary values of three variables of one face cell, loaded ahead of
// use: psi of the stage's input set and, when FULL, the set-A and set-T
// values of the update (by RK kind).
struct AuxVals {
  float p[3], a[3], t[3];
};
__device__ __forceinline__ float sel3(int m, float v0, float v1, float v2) { return m == 0 ? v0 : (m == 1 ? v1 : v2); }

// auxiliary RK update of variable v (abs_RK_beg/inn/fin, mod_abs_npml.F90:738-1105);
// pre: the set-A / set-T values were loaded ahead (pa, pt)
template <int KIND, bool FULL>
__device__ __forceinline__ void aux_rk(const FaceArgs& f, const StageArgs& g, int slot, long li, int v, float psi, float dpsi,
                                       bool pre, float pa, float pt) {
  const long S = 6 * f.cnt, q = v * f.cnt + li;
  float* const d = aux_base(f, slot);
  if constexpr (KIND == 0) {
    d[g.setout * S + q] = psi + g.a * dpsi;
    d[3 * S + q] = psi + g.b * dpsi;
  } else if constexpr (KIND == 1) {
    const float A = FULL && pre ? pa : d[q], T = FULL && pre ? pt : d[3 * S + q];
    d[g.setout * S + q] = A + g.a * dpsi;
    d[3 * S + q] = T + g.b * dpsi;
  } else {
    const float T = FULL && pre ? pt : d[3 * S + q];
    d[g.setout * S + q] = T + g.b * dpsi;
  }
}
__device__ __forceinline__ float aux_in(const FaceArgs& f, const StageArgs& g, int slot, long li, int v) {
  return aux_base(f, slot)[g.setin * 6 * f.cnt + v * f.cnt + li];
}

// Prefetch of the auxiliary values of the first x or y face slab holding a
// cell.  x and y slabs span every plane of the volume, so the face and its
// index step per plane are fixed for the column; the loads of plane k are
// issued with the RK operands, before the plane's barrier.  Cells in more
// than one slab (edges, corners) load the other slabs' values where they are
// applied.
struct AuxPre {
  const float* base;  // set A, first of the three variables, this cell, plane k0
  int face;           // 0..3, or -1: no x or y slab
  int cnt, kst;       // face cell count, index step per plane
  int mask;           // bit n: the cell is in x or y slab n (0..3)
};
__device__ __forceinline__ AuxPre aux_pre(const StageArgs& g, int slot, int i, int j, int k0, int v0, bool comp) {
  AuxPre r{nullptr, -1, 0, 0, 0};
  for (int n = 0; n < 4; ++n) {
    const FaceArgs& f = g.face[n];
    if (comp && in_face(f, i, j, k0)) {
      r.mask |= 1 << n;
      if (r.face < 0) {
        r.face = n;
        r.cnt = (int)f.cnt;
        r.kst = f.ex0 * f.ex1;
        r.base = aux_base(f, slot) + (long)v0 * f.cnt + face_index(f, i, j, k0);
      }
    }
  }
  return r;
}
// slabs holding cell (i, j, k) of a computed row: the column's x / y slabs
// and the plane's z slabs
__device__ __forceinline__ int face_mask(const StageArgs& g, const AuxPre& a, int i, int j, int k) {
  return a.mask | (in_face(g.face[4], i, j, k) ? 16 : 0) | (in_face(g.face[5], i, j, k) ? 32 : 0);
}
template <int KIND, bool FULL>
__device__ __forceinline__ void aux_load(const AuxPre& a, int k, int k0, int setin, AuxVals& r) {
  const float* b = a.base + (long)(k - k0) * a.kst;
  const long c = a.cnt;
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    r.p[m] = b[setin * 6 * c + m * c];
    r.a[m] = FULL && KIND == 1 ? b[m * c] : 0.f;
    r.t[m] = FULL && KIND != 0 ? b[18 * c + m * c] : 0.f;
  }
}
// velocity half of the wide tiles: psi, set A and set T ahead; narrow tiles
// and the stress half (more live RK operands): psi only
constexpr bool AUX_FULL_VEL = true, AUX_FULL_STR = false;

// CFS-PML corrections of hV on the face slabs containing (i, j, k), in the
// reference's face order x1 x2 y1 y2 z1 z2 (mod_abs_npml.F90:1149-2123, the
// stress-derivative half).  D* are the raw one-sided derivatives times the
// 1-D metric (before fluid zeroing).
template <int KIND, bool FULL>
__device__ __forceinline__ void pml_vel(const StageArgs& g, int slot, int i, int j, int k, float h[3], float rr, float DxTxx,
                                        float DxTxy, float DxTxz, float DyTyy, float DyTxy, float DyTyz, float DzTzz,
                                        float DzTxz, float DzTyz, float Txx, float Tyy, float Txy, float Tzz, float Txz,
                                        float Tyz, float rzk, float rxs, float xc, int mask, int pf, const AuxVals pv) {
#pragma nounroll
  for (int n = 0; n < 6; ++n) {
    const FaceArgs& f = g.face[n];
    if (!((mask >> n) & 1)) continue;
    const int ax = n >> 1;
    const int ia = ax == 0 ? i : (ax == 1 ? g.nx + j : g.nx + g.ny + k);
    const float b1 = g.pb1[ia], b1_1 = b1 - 1.0f, ad = g.pad[ia], Dd = g.pdd[ia];
    const long li = face_index(f, i, j, k);
    const bool pre = n == pf;
    const float p3 = pre ? pv.p[0] : aux_in(f, g, slot, li, 3), p4 = pre ? pv.p[1] : aux_in(f, g, slot, li, 4),
                p5 = pre ? pv.p[2] : aux_in(f, g, slot, li, 5);
    float D0, D1, D2;  // corrected derivatives feeding hVx, hVy, hVz
    float q0, q1, q2;  // their psi
    int v0, v1, v2;
    if (ax == 0) {
      D0 = (DxTxx + 2.0f * Txz) * rzk; q0 = p3; v0 = 3;
      D1 = (DxTxy + Tyz) * rzk; q1 = p4; v1 = 4;
      D2 = (DxTxz - Txx + Tzz) * rzk; q2 = p5; v2 = 5;
    } else if (ax == 1) {
      D0 = (DyTxy * rxs + (Txx - Tyy) * xc + Txz) * rzk; q0 = p4; v0 = 4;
      D1 = (DyTyy * rxs + 2.0f * Txy * xc + 2.0f * Tyz) * rzk; q1 = p3; v1 = 3;
      D2 = (DyTyz * rxs + Txz * xc + (Tzz - Tyy)) * rzk; q2 = p5; v2 = 5;
    } else {
      D0 = DzTxz; q0 = p4; v0 = 4;
      D1 = DzTyz; q1 = p5; v1 = 5;
      D2 = DzTzz; q2 = p3; v2 = 3;
    }
    h[0] += (b1_1 * D0 - b1 * q0) * rr;
    h[1] += (b1_1 * D1 - b1 * q1) * rr;
    h[2] += (b1_1 * D2 - b1 * q2) * rr;
    aux_rk<KIND, FULL>(f, g, slot, li, v0, q0, Dd * D0 - ad * q0, pre, sel3(v0 - 3, pv.a[0], pv.a[1], pv.a[2]), sel3(v0 - 3, pv.t[0], pv.t[1], pv.t[2]));
    aux_rk<KIND, FULL>(f, g, slot, li, v1, q1, Dd * D1 - ad * q1, pre, sel3(v1 - 3, pv.a[0], pv.a[1], pv.a[2]), sel3(v1 - 3, pv.t[0], pv.t[1], pv.t[2]));
    aux_rk<KIND, FULL>(f, g, slot, li, v2, q2, Dd * D2 - ad * q2, pre, sel3(v2 - 3, pv.a[0], pv.a[1], pv.a[2]), sel3(v2 - 3, pv.t[0], pv.t[1], pv.t[2]));
  }
}

// CFS-PML corrections of hT (velocity-derivative half), including the
// free-surface Hooke fix-ups of the x and y faces (mod_abs_npml.F90:1303-1333).
// h: Txx Tyy Tzz Txy Txz Tyz.  DzV* are the standard 5-tap z derivatives.
template <int KIND>
__device__ __forceinline__ void pml_str(const StageArgs& g, int slot, int i, int j, int k, float h[6], float lam, float miu,
                                        float DxVx, float DxVy, float DxVz, float DyVx, float DyVy, float DyVz,
                                        float DzVx, float DzVy, float DzVz, float Vx, float Vy, float Vz, float rzk,
                                        float rxs, float xc, int mask, int pf, const AuxVals pv) {
  const float tc13 = lam, tc66 = miu, tc11 = tc13 + 2.0f * tc66;
  const float tc12 = tc13, tc22 = tc11, tc23 = tc13, tc33 = tc11, tc44 = tc66, tc55 = tc66;
  const bool surf = g.freenode && k == g.nk2;
#pragma nounroll
  for (int n = 0; n < 6; ++n) {
    const FaceArgs& f = g.face[n];
    if (!((mask >> n) & 1)) continue;
    const int ax = n >> 1;
    const int ia = ax == 0 ? i : (ax == 1 ? g.nx + j : g.nx + g.ny + k);
    const float b1 = g.pb1[ia], b1_1 = b1 - 1.0f, ad = g.pad[ia], Dd = g.pdd[ia];
    const long li = face_index(f, i, j, k);
    const bool pre = n == pf;
    const float p0 = pre ? pv.p[0] : aux_in(f, g, slot, li, 0), p1 = pre ? pv.p[1] : aux_in(f, g, slot, li, 1),
                p2 = pre ? pv.p[2] : aux_in(f, g, slot, li, 2);
    if (ax == 0) {
      const float D0 = (DxVx + Vz) * rzk, D1 = DxVy * rzk, D2 = (DxVz - Vx) * rzk;
      h[0] += b1_1 * tc11 * D0 - b1 * tc11 * p0;
      h[1] += b1_1 * tc12 * D0 - b1 * tc12 * p0;
      h[2] += b1_1 * tc13 * D0 - b1 * tc13 * p0;
      h[3] += b1_1 * tc66 * D1 - b1 * tc66 * p1;
      h[4] += b1_1 * tc55 * D2 - b1 * tc55 * p2;
      if (surf) {
        h[0] += -b1_1 * tc13 * tc13 / tc33 * D0 + b1 * tc13 * tc13 / tc33 * p0;
        h[1] += -b1_1 * tc23 * tc13 / tc33 * D0 + b1 * tc23 * tc13 / tc33 * p0;
      }
      aux_rk<KIND, AUX_FULL_STR>(f, g, slot, li, 0, p0, Dd * D0 - ad * p0, pre, pv.a[0], pv.t[0]);
      aux_rk<KIND, AUX_FULL_STR>(f, g, slot, li, 1, p1, Dd * D1 - ad * p1, pre, pv.a[1], pv.t[1]);
      aux_rk<KIND, AUX_FULL_STR>(f, g, slot, li, 2, p2, Dd * D2 - ad * p2, pre, pv.a[2], pv.t[2]);
    } else if (ax == 1) {
      const float D0 = (DyVx * rxs - Vy * xc) * rzk;
      const float D1 = (DyVy * rxs + Vx * xc + Vz) * rzk;
      const float D2 = (DyVz * rxs - Vy) * rzk;
      h[3] += b1_1 * tc66 * D0 - b1 * tc66 * p0;
      h[0] += b1_1 * tc12 * D1 - b1 * tc12 * p1;
      h[1] += b1_1 * tc22 * D1 - b1 * tc22 * p1;
      h[2] += b1_1 * tc23 * D1 - b1 * tc23 * p1;
      h[5] += b1_1 * tc44 * D2 - b1 * tc44 * p2;
      if (surf) {
        h[0] += -b1_1 * tc13 * tc23 / tc33 * D1 + b1 * tc13 * tc23 / tc33 * p1;
        h[1] += -b1_1 * tc23 * tc23 / tc33 * D1 + b1 * tc23 * tc23 / tc33 * p1;
      }
      aux_rk<KIND, AUX_FULL_STR>(f, g, slot, li, 0, p0, Dd * D0 - ad * p0, pre, pv.a[0], pv.t[0]);
      aux_rk<KIND, AUX_FULL_STR>(f, g, slot, li, 1, p1, Dd * D1 - ad * p1, pre, pv.a[1], pv.t[1]);
      aux_rk<KIND, AUX_FULL_STR>(f, g, slot, li, 2, p2, Dd * D2 - ad * p2, pre, pv.a[2], pv.t[2]);
    } else {
      h[4] += b1_1 * tc55 * DzVx - b1 * tc55 * p0;
      h[5] += b1_1 * tc44 * DzVy - b1 * tc44 * p1;
      h[0] += b1_1 * tc13 * DzVz - b1 * tc13 * p2;
      h[1] += b1_1 * tc23 * DzVz - b1 * tc23 * p2;
      h[2] += b1_1 * tc33 * DzVz - b1 * tc33 * p2;
      aux_rk<KIND, AUX_FULL_STR>(f, g, slot, li, 0, p0, Dd * DzVx - ad * p0, pre, pv.a[0], pv.t[0]);
      aux_rk<KIND, AUX_FULL_STR>(f, g, slot, li, 1, p1, Dd * DzVy - ad * p1, pre, pv.a[1], pv.t[1]);
      aux_rk<KIND, AUX_FULL_STR>(f, g, slot, li, 2, p2, Dd * DzVz - ad * p2, pre, pv.a[2], pv.t[2]);
    }
  }
}

// one-sided 5-tap operator on LDS: p points at the centre, s is the stride
template <int FWD>
__device__ __forceinline__ float dlds(const float* p, int s) {
  if constexpr (FWD)
    return MCA_1 * p[-s] + MCA0 * p[0] + MCA1 * p[s] + MCA2 * p[2 * s] + MCA3 * p[3 * s];
  else
    return -MCA3 * p[-3 * s] - MCA2 * p[-2 * s] - MCA1 * p[-s] - MCA0 * p[0] - MCA_1 * p[s];
}
// z window of 5 slots: F holds offsets -1..+3, B holds -3..+1
template <int FWD>
__device__ __forceinline__ float dz5(const float* w) {
  if constexpr (FWD)
    return MCA_1 * w[0] + MCA0 * w[1] + MCA1 * w[2] + MCA2 * w[3] + MCA3 * w[4];
  else
    return -MCA3 * w[0] - MCA2 * w[1] - MCA1 * w[2] - MCA0 * w[3] - MCA_1 * w[4];
}
template <int FWD>
__device__ __forceinline__ float dz24(const float* w) {
  if constexpr (FWD)
    return M24A0 * w[1] + M24A1 * w[2] + M24A2 * w[3];
  else
    return -M24A2 * w[1] - M24A1 * w[2] - M24A0 * w[3];
}
template <int FWD>
__device__ __forceinline__ float dz22(const float* w) {
  if constexpr (FWD)
    return -w[1] + w[2];
  else
    return -w[2] + w[3];
}
template <int FWD>
__device__ __forceinline__ float zc(const float* w) {
  return FWD ? w[1] : w[3];
}

// Halo slot of one thread in one LDS plane: x halo (XL left + XR right
// columns of the TY rows, threads tid < 4*TY) or y halo (YL + YR rows of the
// TW columns, every thread).  Slots outside the volume or beyond what the
// tile's computed columns ie / rows je read (partial tiles) hold zero and
// never feed a computed cell.
struct HaloSlot {
  unsigned off;
  int lds;
  bool ok;
};
template <int XF, int TW, int TY, int LW>
__device__ __forceinline__ HaloSlot xhalo_slot(const StageArgs& g, int i0, int j0, int ie, int je, int tid) {
  constexpr int XL = XF ? 1 : 3;
  HaloSlot h{0u, 0, false};
  if (tid < 4 * TY) {
    const int r = tid % TY, s = tid / TY;
    const int col = s < XL ? i0 - XL + s : i0 + TW + (s - XL);
    const int row = j0 + r;
    h.ok = col < g.nx && col <= ie + 3 && row < g.ny && row <= je;
    h.off = h.ok ? (unsigned)(col + g.sy * row) : 0u;
    h.lds = (r + 3) * LW + (col - i0 + 3);
  } else {
    h.lds = -1;
  }
  return h;
}
template <int YF, int TW, int TY, int LW>
__device__ __forceinline__ HaloSlot yhalo_slot(const StageArgs& g, int i0, int j0, int ie, int je, int tid) {
  constexpr int YL = YF ? 1 : 3;
  HaloSlot h{0u, -1, false};
  if (tid >= 4 * TW) return h;
  const int c = tid % TW, s = tid / TW;
  const int row = s < YL ? j0 - YL + s : j0 + TY + (s - YL);
  const int col = i0 + c;
  h.ok = col < g.nx && col <= ie && row < g.ny && row <= je + 3;
  h.off = h.ok ? (unsigned)(col + g.sy * row) : 0u;
  h.lds = (row - j0 + 3) * LW + (c + 3);
  return h;
}
__device__ __forceinline__ float ld(const float* __restrict__ p, const HaloSlot& h, unsigned kz) {
  return h.ok ? p[h.off + kz] : 0.f;
}


// ---------------------------------------------------------------------------
// k_vel: momentum equations with the spherical geometric terms
// (mod_macdrp.F90:928-933), fluid zeroing of shear-stress derivatives
// (:917-926), RK update of Vx Vy Vz.
// LDS: Txx (x), Txy (x, y), Txz (x), Tyy (y), Tyz (y); z window: Tzz Txz Tyz.
// ---------------------------------------------------------------------------
template <int BYT, int RY, int TW>
struct VelLds {
  static constexpr int TY = BYT * (TX / TW) * RY, LW = TW + 6, LH = TY + 6, PL = LW * LH, NL = 5, SIZE = 2 * NL * PL;
};
template <int XF, int YF, int ZF, int KIND, int BYT, int RY, int TW, bool PML>
__device__ __forceinline__ void vel_body(const StageArgs& g, const int4 it, int slot, float* shbase) {
  // RS: rows per row group (the stride between a thread's rows)
  constexpr int RS = BYT * (TX / TW), TY = RS * RY, LW = TW + 6, LH = TY + 6, PL = LW * LH, NL = 5;
  // set-A / set-T values of the CFS-PML auxiliaries loaded ahead in wide tiles
  constexpr bool FULLV = AUX_FULL_VEL && TW == TX;
  float(*sh)[NL * PL] = reinterpret_cast<float(*)[NL * PL]>(shbase);  // Txx Txy Txz Tyy Tyz, double buffered
  const int tx = threadIdx.x % TW, ty = threadIdx.x / TW + (TX / TW) * threadIdx.y, tid = threadIdx.x + TX * threadIdx.y;
  const int i0 = it.x & 0xffff, j0 = it.y & 0xffff, k0 = it.z, k1 = it.w;
  // last computed column and row: tiles end at CFS-PML slab boundaries
  const int ie = i0 + (it.x >> 16) - 1, je = j0 + (it.y >> 16) - 1;
  const bool tile_pml = PML && ((g.face[0].on && i0 <= g.face[0].hi0) || (g.face[1].on && ie >= g.face[1].lo0) ||
                                (g.face[2].on && j0 <= g.face[2].hi1) || (g.face[3].on && je >= g.face[3].lo1));
  const long fs = g.fs;
  const unsigned sz = (unsigned)g.sz;
  const long bo = (long)slot * g.bstride;
  const float* __restrict__ W = g.win + bo;
  float* __restrict__ wout = g.wout + bo;
  const float* __restrict__ mm = g.m + bo;
  float* __restrict__ tt = g.t + bo;
  const float* __restrict__ pTxx = W + TXX * fs;
  const float* __restrict__ pTyy = W + TYY * fs;
  const float* __restrict__ pTxy = W + TXY * fs;
  const float* __restrict__ pTzz = W + TZZ * fs;
  const float* __restrict__ pTxz = W + TXZ * fs;
  const float* __restrict__ pTyz = W + TYZ * fs;
  const int i = i0 + tx;
  const bool iok = i < g.nx && i <= ie + 3, icomp = i <= ie;
  const int ic = min(i, g.nx - 1);
  const float xi = g.xi_x[ic], rxs = g.rxsin[ic], xc = g.xcot[ic];
  const HaloSlot hx = xhalo_slot<XF, TW, TY, LW>(g, i0, j0, ie, je, tid);
  const HaloSlot hy = yhalo_slot<YF, TW, TY, LW>(g, i0, j0, ie, je, tid);

  float w[RY][3][5];  // Tzz Txz Tyz
  float cTxx[RY], cTyy[RY], cTxy[RY];
  bool ok[RY], comp[RY], fluid[RY];
  unsigned cc[RY];
  float et[RY];
#pragma unroll
  for (int q = 0; q < RY; ++q) {
    const int j = j0 + ty + RS * q;
    ok[q] = iok && j < g.ny && j <= je + 3;
    comp[q] = icomp && j <= je;
    cc[q] = ok[q] ? (unsigned)(i + g.sy * j) : 0u;
    et[q] = g.eta_y[min(j, g.ny - 1)];
    fluid[q] = comp[q] ? g.fluidcol[cc[q]] != 0 : false;
    const unsigned c0 = cc[q] + sz * (unsigned)k0;
#pragma unroll
    for (int s = 0; s < 5; ++s) {
      const int o = ZF ? s - 1 : s - 3;
      w[q][0][s] = ok[q] ? pTzz[c0 + o * (int)sz] : 0.f;
      w[q][1][s] = ok[q] ? pTxz[c0 + o * (int)sz] : 0.f;
      w[q][2][s] = ok[q] ? pTyz[c0 + o * (int)sz] : 0.f;
    }
    cTxx[q] = ok[q] ? pTxx[c0] : 0.f;
    cTyy[q] = ok[q] ? pTyy[c0] : 0.f;
    cTxy[q] = ok[q] ? pTxy[c0] : 0.f;
  }
  AuxPre pre[RY];
#pragma unroll
  for (int q = 0; q < RY; ++q) pre[q] = PML ? aux_pre(g, slot, i, j0 + ty + RS * q, k0, 3, comp[q]) : AuxPre{nullptr, -1, 0, 0};
  unsigned kz = sz * (unsigned)k0;
  float hxTxx = ld(pTxx, hx, kz), hxTxy = ld(pTxy, hx, kz), hxTxz = ld(pTxz, hx, kz);
  float hyTxy = ld(pTxy, hy, kz), hyTyy = ld(pTyy, hy, kz), hyTyz = ld(pTyz, hy, kz);

  for (int k = k0; k <= k1; ++k, kz += sz) {
    float* s = sh[(k - k0) & 1];
    float* sTxx = s;
    float* sTxy = s + PL;
    float* sTxz = s + 2 * PL;
    float* sTyy = s + 3 * PL;
    float* sTyz = s + 4 * PL;
#pragma unroll
    for (int q = 0; q < RY; ++q) {
      const int o = (ty + RS * q + 3) * LW + tx + 3;
      sTxx[o] = cTxx[q];
      sTxy[o] = cTxy[q];
      sTyy[o] = cTyy[q];
      sTxz[o] = zc<ZF>(w[q][1]);
      sTyz[o] = zc<ZF>(w[q][2]);
    }
    if (hx.lds >= 0) {
      sTxx[hx.lds] = hxTxx;
      sTxy[hx.lds] = hxTxy;
      sTxz[hx.lds] = hxTxz;
    }
    if (hy.lds >= 0) {
      sTxy[hy.lds] = hyTxy;
      sTyy[hy.lds] = hyTyy;
      sTyz[hy.lds] = hyTyz;
    }
    // RK operands of plane k
    float rr[RY], mo[RY][3], to[RY][3];
    bool fl[RY];
#pragma unroll
    for (int q = 0; q < RY; ++q) {
      const unsigned c = cc[q] + kz;
      rr[q] = comp[q] ? lds_nt(g.rrho + c) : 0.f;
      fl[q] = fluid[q] ? g.mu[c] <= SEIS_ZERO : false;
#pragma unroll
      for (int f = 0; f < 3; ++f) {
        const long e = f * fs + c;
        mo[q][f] = (comp[q] && KIND != 2) ? lds_nt(KIND == 0 ? W + e : mm + e) : 0.f;
        to[q][f] = (comp[q] && KIND != 0) ? lds_nt(tt + e) : 0.f;
      }
    }
    // prefetch plane k+1
    const bool more = k < k1;
    float nTxx[RY], nTyy[RY], nTxy[RY], lead[RY][3];
    const unsigned kn = kz + sz;
#pragma unroll
    for (int q = 0; q < RY; ++q) {
      const unsigned c = cc[q] + kn;
      const unsigned l = c + (ZF ? 3u : 1u) * sz;
      const bool p = more && ok[q];
      lead[q][0] = p ? pTzz[l] : 0.f;
      lead[q][1] = p ? pTxz[l] : 0.f;
      lead[q][2] = p ? pTyz[l] : 0.f;
      nTxx[q] = p ? pTxx[c] : 0.f;
      nTyy[q] = p ? pTyy[c] : 0.f;
      nTxy[q] = p ? pTxy[c] : 0.f;
    }
    float nhxTxx = 0.f, nhxTxy = 0.f, nhxTxz = 0.f, nhyTxy = 0.f, nhyTyy = 0.f, nhyTyz = 0.f;
    if (more) {
      nhxTxx = ld(pTxx, hx, kn); nhxTxy = ld(pTxy, hx, kn); nhxTxz = ld(pTxz, hx, kn);
      nhyTxy = ld(pTxy, hy, kn); nhyTyy = ld(pTyy, hy, kn); nhyTyz = ld(pTyz, hy, kn);
    }
    AuxVals pv[RY];
    if constexpr (PML) {
#pragma unroll
      for (int q = 0; q < RY; ++q)
        if (pre[q].face >= 0) aux_load<KIND, FULLV>(pre[q], k, k0, g.setin, pv[q]);
    }
    __syncthreads();
    const float ze = g.zeta_z[k], rzk = g.rz[k];
    const bool pml_here = PML && (tile_pml || (g.face[4].on && k <= g.face[4].hi2) || (g.face[5].on && k >= g.face[5].lo2));
#pragma unroll
    for (int q = 0; q < RY; ++q) {
      if (comp[q]) {
        const unsigned c = cc[q] + kz;
        const int o = (ty + RS * q + 3) * LW + tx + 3;
        const float DxTxx = dlds<XF>(sTxx + o, 1) * xi;
        float DxTxy = dlds<XF>(sTxy + o, 1) * xi;
        float DxTxz = dlds<XF>(sTxz + o, 1) * xi;
        const float DyTyy = dlds<YF>(sTyy + o, LW) * et[q];
        float DyTxy = dlds<YF>(sTxy + o, LW) * et[q];
        float DyTyz = dlds<YF>(sTyz + o, LW) * et[q];
        const float DzTzz = dz5<ZF>(w[q][0]) * ze;
        float DzTxz = dz5<ZF>(w[q][1]) * ze;
        float DzTyz = dz5<ZF>(w[q][2]) * ze;
        const float rDxTxy = DxTxy, rDxTxz = DxTxz, rDyTxy = DyTxy, rDyTyz = DyTyz, rDzTxz = DzTxz, rDzTyz = DzTyz;
        if (fl[q]) {
          DxTxy = 0.f; DxTxz = 0.f; DyTxy = 0.f; DyTyz = 0.f; DzTxz = 0.f; DzTyz = 0.f;
        }
        const float Txx = cTxx[q], Tyy = cTyy[q], Txy = cTxy[q];
        const float Tzz = zc<ZF>(w[q][0]), Txz = zc<ZF>(w[q][1]), Tyz = zc<ZF>(w[q][2]);
        float h[3];
        h[0] = rr[q] * (DxTxx * rzk + DyTxy * rzk * rxs + DzTxz + (3.0f * Txz + Txx * xc - Tyy * xc) * rzk);
        h[1] = rr[q] * (DxTxy * rzk + DyTyy * rzk * rxs + DzTyz + (2.0f * Txy * xc + 3.0f * Tyz) * rzk);
        h[2] = rr[q] * (DxTxz * rzk + DyTyz * rzk * rxs + DzTzz + (2.0f * Tzz - Txx - Tyy + Txz * xc) * rzk);
        if constexpr (PML)
          if (pml_here) pml_vel<KIND, FULLV>(g, slot, i, j0 + ty + RS * q, k, h, rr[q], DxTxx, rDxTxy, rDxTxz, DyTyy, rDyTxy, rDyTyz, DzTzz,
                        rDzTxz, rDzTyz, Txx, Tyy, Txy, Tzz, Txz, Tyz, rzk, rxs, xc, face_mask(g, pre[q], i, j0 + ty + RS * q, k), pre[q].face, pv[q]);
#pragma unroll
        for (int f = 0; f < 3; ++f) {
          const long e = f * fs + c;
          if constexpr (KIND == 2) {
            st_nt(wout + e, to[q][f] + g.b * h[f]);
          } else {
            st_nt(wout + e, mo[q][f] + g.a * h[f]);
            st_nt(tt + e, (KIND == 0 ? mo[q][f] : to[q][f]) + g.b * h[f]);
          }
        }
      }
      // advance the window and the plane registers
#pragma unroll
      for (int f = 0; f < 3; ++f) {
#pragma unroll
        for (int sl = 0; sl < 4; ++sl) w[q][f][sl] = w[q][f][sl + 1];
        w[q][f][4] = lead[q][f];
      }
      cTxx[q] = nTxx[q];
      cTyy[q] = nTyy[q];
      cTxy[q] = nTxy[q];
    }
    hxTxx = nhxTxx; hxTxy = nhxTxy; hxTxz = nhxTxz;
    hyTxy = nhyTxy; hyTyy = nhyTyy; hyTyz = nhyTyz;
  }
}

// ---------------------------------------------------------------------------
// k_str: strain rates (mod_macdrp.F90:935-941), near-surface low-order z
// operators (CondFreeVLOW, :872-915), the traction condition at k = nk2
// with zero source terms (:942-950; k_src adds the source part), Hooke
// (:952-957), RK update of the six stresses.  At k = nk2 the RK "m" values of
// Tzz Txz Tyz are the pre-image stresses saved by k_surface.
// LDS: Vx Vy Vz (x and y halos); z window: Vx Vy Vz.
// ---------------------------------------------------------------------------
template <int BYT, int RY, int TW>
struct StrLds {
  static constexpr int TY = BYT * (TX / TW) * RY, LW = TW + 6, LH = TY + 6, PL = LW * LH, SIZE = 2 * 3 * PL;
};
template <int XF, int YF, int ZF, int KIND, int BYT, int RY, int TW, bool PML>
__device__ __forceinline__ void str_body(const StageArgs& g, const int4 it, int slot, float* shbase) {
  constexpr int RS = BYT * (TX / TW), TY = RS * RY, LW = TW + 6, LH = TY + 6, PL = LW * LH;
  float(*sh)[3 * PL] = reinterpret_cast<float(*)[3 * PL]>(shbase);  // Vx Vy Vz, double buffered
  const int tx = threadIdx.x % TW, ty = threadIdx.x / TW + (TX / TW) * threadIdx.y, tid = threadIdx.x + TX * threadIdx.y;
  const int i0 = it.x & 0xffff, j0 = it.y & 0xffff, k0 = it.z, k1 = it.w;
  // last computed column and row: tiles end at CFS-PML slab boundaries
  const int ie = i0 + (it.x >> 16) - 1, je = j0 + (it.y >> 16) - 1;
  const bool tile_pml = PML && ((g.face[0].on && i0 <= g.face[0].hi0) || (g.face[1].on && ie >= g.face[1].lo0) ||
                                (g.face[2].on && j0 <= g.face[2].hi1) || (g.face[3].on && je >= g.face[3].lo1));
  const long fs = g.fs;
  const unsigned sz = (unsigned)g.sz;
  const long bo = (long)slot * g.bstride;
  const float* __restrict__ W = g.win + bo;
  float* __restrict__ wout = g.wout + bo;
  const float* __restrict__ mm = g.m + bo;
  float* __restrict__ tt = g.t + bo;
  const float* __restrict__ pV0 = W + VX * fs;
  const float* __restrict__ pV1 = W + VY * fs;
  const float* __restrict__ pV2 = W + VZ * fs;
  const int i = i0 + tx;
  const bool iok = i < g.nx && i <= ie + 3, icomp = i <= ie;
  const int ic = min(i, g.nx - 1);
  const float xi = g.xi_x[ic], rxs = g.rxsin[ic], xc = g.xcot[ic];
  const HaloSlot hx = xhalo_slot<XF, TW, TY, LW>(g, i0, j0, ie, je, tid);
  const HaloSlot hy = yhalo_slot<YF, TW, TY, LW>(g, i0, j0, ie, je, tid);

  float w[RY][3][5];
  bool ok[RY], comp[RY];
  unsigned cc[RY];
  float et[RY];
#pragma unroll
  for (int q = 0; q < RY; ++q) {
    const int j = j0 + ty + RS * q;
    ok[q] = iok && j < g.ny && j <= je + 3;
    comp[q] = icomp && j <= je;
    cc[q] = ok[q] ? (unsigned)(i + g.sy * j) : 0u;
    et[q] = g.eta_y[min(j, g.ny - 1)];
    const unsigned c0 = cc[q] + sz * (unsigned)k0;
#pragma unroll
    for (int s = 0; s < 5; ++s) {
      const int o = ZF ? s - 1 : s - 3;
      w[q][0][s] = ok[q] ? pV0[c0 + o * (int)sz] : 0.f;
      w[q][1][s] = ok[q] ? pV1[c0 + o * (int)sz] : 0.f;
      w[q][2][s] = ok[q] ? pV2[c0 + o * (int)sz] : 0.f;
    }
  }
  AuxPre pre[RY];
#pragma unroll
  for (int q = 0; q < RY; ++q) pre[q] = PML ? aux_pre(g, slot, i, j0 + ty + RS * q, k0, 0, comp[q]) : AuxPre{nullptr, -1, 0, 0};
  unsigned kz = sz * (unsigned)k0;
  float hx0 = ld(pV0, hx, kz), hx1 = ld(pV1, hx, kz), hx2 = ld(pV2, hx, kz);
  float hy0 = ld(pV0, hy, kz), hy1 = ld(pV1, hy, kz), hy2 = ld(pV2, hy, kz);

  for (int k = k0; k <= k1; ++k, kz += sz) {
    float* s = sh[(k - k0) & 1];
#pragma unroll
    for (int q = 0; q < RY; ++q) {
      const int o = (ty + RS * q + 3) * LW + tx + 3;
      s[o] = zc<ZF>(w[q][0]);
      s[PL + o] = zc<ZF>(w[q][1]);
      s[2 * PL + o] = zc<ZF>(w[q][2]);
    }
    if (hx.lds >= 0) {
      s[hx.lds] = hx0;
      s[PL + hx.lds] = hx1;
      s[2 * PL + hx.lds] = hx2;
    }
    if (hy.lds >= 0) {
      s[hy.lds] = hy0;
      s[PL + hy.lds] = hy1;
      s[2 * PL + hy.lds] = hy2;
    }
    const bool surf = g.freenode && k == g.nk2;
    // RK operands of plane k
    float la[RY], mu_[RY], mo[RY][6], to[RY][6];
#pragma unroll
    for (int q = 0; q < RY; ++q) {
      const unsigned c = cc[q] + kz;
      la[q] = comp[q] ? lds_nt(g.lam + c) : 0.f;
      mu_[q] = comp[q] ? lds_nt(g.mu + c) : 0.f;
#pragma unroll
      for (int f = 0; f < 6; ++f) {
        const long e = (TXX + f) * fs + c;
        float mv = (comp[q] && KIND != 2) ? lds_nt(KIND == 0 ? W + e : mm + e) : 0.f;
        if (surf && comp[q] && KIND != 2 && (f == 2 || f == 4 || f == 5))
          mv = g.msurf[(long)slot * 3 * g.ps + (f == 2 ? 0 : (f == 4 ? 1 : 2)) * g.ps + cc[q]];
        mo[q][f] = mv;
        to[q][f] = (comp[q] && KIND != 0) ? lds_nt(tt + e) : 0.f;
      }
    }
    // prefetch plane k+1
    const bool more = k < k1;
    const unsigned kn = kz + sz;
    float lead[RY][3];
#pragma unroll
    for (int q = 0; q < RY; ++q) {
      const unsigned l = cc[q] + kn + (ZF ? 3u : 1u) * sz;
      const bool p = more && ok[q];
      lead[q][0] = p ? pV0[l] : 0.f;
      lead[q][1] = p ? pV1[l] : 0.f;
      lead[q][2] = p ? pV2[l] : 0.f;
    }
    float nx0 = 0.f, nx1 = 0.f, nx2 = 0.f, ny0 = 0.f, ny1 = 0.f, ny2 = 0.f;
    if (more) {
      nx0 = ld(pV0, hx, kn); nx1 = ld(pV1, hx, kn); nx2 = ld(pV2, hx, kn);
      ny0 = ld(pV0, hy, kn); ny1 = ld(pV1, hy, kn); ny2 = ld(pV2, hy, kn);
    }
    AuxVals pv[RY];
    if constexpr (PML) {
#pragma unroll
      for (int q = 0; q < RY; ++q)
        if (pre[q].face >= 0) aux_load<KIND, AUX_FULL_STR>(pre[q], k, k0, g.setin, pv[q]);
    }
    __syncthreads();
    const float ze = g.zeta_z[k], rzk = g.rz[k];
    const bool pml_here = PML && (tile_pml || (g.face[4].on && k <= g.face[4].hi2) || (g.face[5].on && k >= g.face[5].lo2));
#pragma unroll
    for (int q = 0; q < RY; ++q) {
      if (comp[q]) {
        const unsigned c = cc[q] + kz;
        const int o = (ty + RS * q + 3) * LW + tx + 3;
        const float DxVx = dlds<XF>(s + o, 1) * xi;
        const float DxVy = dlds<XF>(s + PL + o, 1) * xi;
        const float DxVz = dlds<XF>(s + 2 * PL + o, 1) * xi;
        const float DyVx = dlds<YF>(s + o, LW) * et[q];
        const float DyVy = dlds<YF>(s + PL + o, LW) * et[q];
        const float DyVz = dlds<YF>(s + 2 * PL + o, LW) * et[q];
        const float sDzVx = dz5<ZF>(w[q][0]) * ze, sDzVy = dz5<ZF>(w[q][1]) * ze, sDzVz = dz5<ZF>(w[q][2]) * ze;
        float DzVx = sDzVx, DzVy = sDzVy, DzVz = sDzVz;
        if (g.freenode && k == g.nk2 - 1) {
          DzVx = dz22<ZF>(w[q][0]) * ze; DzVy = dz22<ZF>(w[q][1]) * ze; DzVz = dz22<ZF>(w[q][2]) * ze;
        } else if (g.freenode && k == g.nk2 - 2) {
          DzVx = dz24<ZF>(w[q][0]) * ze; DzVy = dz24<ZF>(w[q][1]) * ze; DzVz = dz24<ZF>(w[q][2]) * ze;
        }
        const float Vx = zc<ZF>(w[q][0]), Vy = zc<ZF>(w[q][1]), Vz = zc<ZF>(w[q][2]);
        const float lam = la[q], miu = mu_[q], lam2mu = lam + 2.0f * miu;
        const float E11 = (DxVx + Vz) * rzk;
        const float E22 = (Vx * xc + DyVy * rxs + Vz) * rzk;
        const float E12 = (DxVy + DyVx * rxs - Vy * xc) * rzk * 0.5f;
        float E33 = DzVz;
        float E13 = (DxVz * rzk + DzVx - Vx * rzk) * 0.5f;
        float E23 = (DyVz * rzk * rxs + DzVy - Vy * rzk) * 0.5f;
        if (surf) {
          E33 = -lam / lam2mu * (E11 + E22);
          E13 = 0.f;
          E23 = 0.f;
        }
        float h[6];  // Txx Tyy Tzz Txy Txz Tyz
        h[0] = lam2mu * E11 + lam * (E22 + E33);
        h[1] = lam2mu * E22 + lam * (E11 + E33);
        h[2] = lam2mu * E33 + lam * (E11 + E22);
        h[3] = 2.0f * miu * E12;
        h[4] = 2.0f * miu * E13;
        h[5] = 2.0f * miu * E23;
        if constexpr (PML)
          if (pml_here) pml_str<KIND>(g, slot, i, j0 + ty + RS * q, k, h, lam, miu, DxVx, DxVy, DxVz, DyVx, DyVy, DyVz, sDzVx, sDzVy,
                        sDzVz, Vx, Vy, Vz, rzk, rxs, xc, face_mask(g, pre[q], i, j0 + ty + RS * q, k), pre[q].face, pv[q]);
#pragma unroll
        for (int f = 0; f < 6; ++f) {
          const long e = (TXX + f) * fs + c;
          if constexpr (KIND == 2) {
            st_nt(wout + e, to[q][f] + g.b * h[f]);
          } else {
            st_nt(wout + e, mo[q][f] + g.a * h[f]);
            st_nt(tt + e, (KIND == 0 ? mo[q][f] : to[q][f]) + g.b * h[f]);
          }
        }
      }
#pragma unroll
      for (int f = 0; f < 3; ++f) {
#pragma unroll
        for (int sl = 0; sl < 4; ++sl) w[q][f][sl] = w[q][f][sl + 1];
        w[q][f][4] = lead[q][f];
      }
    }
    hx0 = nx0; hx1 = nx1; hx2 = nx2;
    hy0 = ny0; hy1 = ny1; hy2 = ny2;
  }
}

// One launch per stage and kind: the narrow edge tiles run the narrow
// body (with the CFS-PML corrections), the wide PML-shell tiles the PML
// body, the wide interior tiles the lean body.  At most 168 VGPRs (3 waves
// per SIMD): the inner RK kind of k_vel then keeps a few values in scratch,
// which measures faster than 2 waves per SIMD.  With 4-row tiles the first
// and last RK kinds of k_vel fit 96 VGPRs (5 waves per SIMD; 0.4% faster
// per stage, the inner kind and k_str spill too much there).
template <int XF, int YF, int ZF, int KIND, int BYT, int RY>
__global__ __launch_bounds__(TX* BYT) __attribute__((amdgpu_waves_per_eu(BYT * RY == 4 && KIND != 1 ? 5 : 3))) void k_vel(StageArgs g) {
  constexpr int SW = VelLds<BYT, RY, TX>::SIZE, SN = VelLds<BYT, RY, TXN>::SIZE;
  __shared__ float sh[SW > SN ? SW : SN];
  const ItemPos ip = item_pos(g);
  const int L = ip.item, slot = g.slot0 + ip.member;
  const int4 it = g.items[L];
  if (L < g.p1)
    vel_body<XF, YF, ZF, KIND, BYT, RY, TXN, true>(g, it, slot, sh);
  else if (L < g.p2)
    vel_body<XF, YF, ZF, KIND, BYT, RY, TX, true>(g, it, slot, sh);
  else
    vel_body<XF, YF, ZF, KIND, BYT, RY, TX, false>(g, it, slot, sh);
}
template <int XF, int YF, int ZF, int KIND, int BYT, int RY>
__global__ __launch_bounds__(TX* BYT) __attribute__((amdgpu_waves_per_eu(3))) void k_str(StageArgs g) {
  constexpr int SW = StrLds<BYT, RY, TX>::SIZE, SN = StrLds<BYT, RY, TXN>::SIZE;
  __shared__ float sh[SW > SN ? SW : SN];
  const ItemPos ip = item_pos(g);
  const int L = ip.item, slot = g.slot0 + ip.member;
  const int4 it = g.items[L];
  if (L < g.p1)
    str_body<XF, YF, ZF, KIND, BYT, RY, TXN, true>(g, it, slot, sh);
  else if (L < g.p2)
    str_body<XF, YF, ZF, KIND, BYT, RY, TX, true>(g, it, slot, sh);
  else
    str_body<XF, YF, ZF, KIND, BYT, RY, TX, false>(g, it, slot, sh);
}

// Stage-kernel tile shapes (the autotuner picks among them; the first is the
// default): 4-row tiles of 4 waves with 1 row per thread, 8-row tiles of 4
// waves with 2 rows per thread, 16-row tiles of 8 waves, 8-row tiles of 8
// waves with 1 row per thread.  At 400x400x200 with PML: 1250, 1293, 1481,
// 1259 us per stage.
constexpr int NTILECFG = 4;
constexpr int TILE_BYT[NTILECFG] = {4, 4, 8, 8};  // waves per block
constexpr int TILE_RY[NTILECFG] = {1, 2, 2, 1};   // rows per thread

// The 8 F/B patterns x 3 RK kinds of one stage kernel (0: k_vel, 1: k_str)
// and tile shape are instantiated in a translation unit of their own
// (fw_stage_inst.hip, compiled once per kernel and shape, in parallel).
typedef void (*StageFn)(StageArgs);
StageFn stage_fn_v0(int pat, int kind);
StageFn stage_fn_v1(int pat, int kind);
StageFn stage_fn_v2(int pat, int kind);
StageFn stage_fn_v3(int pat, int kind);
StageFn stage_fn_s0(int pat, int kind);
StageFn stage_fn_s1(int pat, int kind);
StageFn stage_fn_s2(int pat, int kind);
StageFn stage_fn_s3(int pat, int kind);

}  // namespace fw
