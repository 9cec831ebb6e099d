// Stage kernel of libfwant: the operator sweep Lx?_Ly?_Lz?
// (code/srcF/mod_macdrp.F90:790-962) fused with the RK update
// (mod_macdrp.F90:291-393) and the CFS-PML corrections
// (mod_abs_npml.F90:1149-2123, 738-1105), one launch per RK stage.
//
// A block holds two wave groups on the same tile and planes:
//   velocity group (waves 0 .. BYT-1)   hVx hVy hVz from the six stresses
//   stress group   (waves BYT .. 2BYT-1) hTxx .. hTyz from the three velocities
// The RHS of a velocity depends on stresses only and vice versa, so each
// group keeps only its own stencil window in registers; the current plane of
// each group's stencil fields is staged in LDS, and each group reads the
// other group's plane centres from there: the stage input's own values at a
// cell (the RK "m" of the first stage, and W_1 / W_3 of the RK form below)
// cost no HBM reads.
//
// RK form.  The reference runs classical RK4 in a 4-register form (m = W_n;
// stage s evaluates h_s on W_s; W_{s+1} = m + a_s h_s, t accumulates
// b_s h_s; W_{n+1} = t + b_3 h_3).  Since b_0 = a_0 / 3 and b_2 = a_2 / 3
// (a = 1/2, 1/2, 1; b = 1/6, 1/3, 1/3, 1/6), the accumulated b_0 h_0 and
// b_2 h_2 equal (W_1 - m) / 3 and (W_3 - m) / 3, and the accumulator is
// written at stage 1 only, with the -m / 3 of stage 3 folded in, so that
// the last stage needs no m:
//   form 0: W_1 = m + a_0 h_0                         (m = W_n = stage input)
//   form 1: W_2 = m + a_1 h_1,  t = c m + r W_1 + b_1 h_1   (c = 1 - 2r)
//   form 2: W_3 = m + a_2 h_2
//   form 3: W_{n+1} = t + r W_3 + b_3 h_3               (r = 1/3)
// which moves 120 floats per cell and step instead of 156 (the same form
// runs on the CFS-PML auxiliaries).  Source terms added after a stage
// (k_src) add a_s dh to W_{s+1} only at forms 0 and 2, so they enter t
// through the same differences.
//
// Work decomposition: a block owns a TW x TY tile of (i, j) columns over a
// chunk of k planes and marches k.  Tiles are TW = 64 columns wide (a wave
// row is one 256-byte row segment), except the narrow tiles at the x edge
// of the volume (the remainder of the x extent after whole 64-column tiles,
// when it is at most 32 columns), which are TW = 16 columns wide: a wave then
// covers 4 rows of 16 columns, and no lanes idle on a 16-column remainder.
// Planes in LDS carry their one-sided halo (F: -1..+3, B: -3..+1), double
// buffered so one barrier per plane suffices (both groups meet at the same
// barrier); z derivatives come from a 5-slot register window per column.
// While plane k is computed, the loads of plane k+1 (window lead, halo) and
// the RK operands of plane k are already in flight.  In a 64-wide tile wave
// w of a group owns rows j0 + w + BYT q (q < RY): every global access of a
// wave is 64 consecutive floats of one row.  Offsets are 32-bit relative to
// per-field base pointers.
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
// RK operands (m, t, media) of plane k+1 loaded while plane k is computed
#ifndef FW_RK_AHEAD
#define FW_RK_AHEAD 0
#endif
constexpr bool RK_AHEAD = FW_RK_AHEAD != 0;

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
  const float* __restrict__ win;  // stage input W_s
  float* __restrict__ wout;       // stage output (set A at form 3, where m is read through wout)
  const float* __restrict__ m;    // set A = W_n
  float* __restrict__ t;          // accumulator
  const float* __restrict__ lam;
  const float* __restrict__ mu;
  const float* __restrict__ rrho;
  const float* __restrict__ xi_x;
  const float* __restrict__ eta_y;
  const float* __restrict__ zeta_z;
  const float* __restrict__ rz;
  const float* __restrict__ rxsin;
  const float* __restrict__ xcot;
  // surface copies at k = nk2 (Tzz, Txz, Tyz), 6 planes per member: 0-2 the
  // pre-image stage-0 input (the "m" there), 3-5 the pre-image input of forms
  // 1 and 3 (the W_1 / W_3 there)
  const float* __restrict__ msurf;
  const int* __restrict__ fluidcol;      // 1 where the (i, j) column has a cell with mu <= SEIS_ZERO
  long fs, sy, sz, ps;
  int nx, ny;
  int ni1, ni2, nj1, nj2, nk1, nk2;
  int freenode;
  const int4* __restrict__ items;  // work items (i0 | ncol << 16, j0 | nrow << 16, k0, k1): a tile over a chunk of planes
  int nitems, p1, p2;  // item parts: narrow [0, p1), wide PML [p1, p2), wide interior
  float a, b, r;       // a_s dt, b_s dt, r = b_{s-1} / a_{s-1}
  float c;             // form 1: the share of m kept in the accumulator, 1 - r - b_2 / a_2
  // CFS-PML: per-axis profiles 1/B, A + D/B, D/B (x: nx, y: ny, z: nz entries)
  const float* __restrict__ pb1;
  const float* __restrict__ pad;
  const float* __restrict__ pdd;
  int n1, setin, setout, pml;
  FaceArgs face[6];
  // batched virtual sources: nm members from slot0 in one launch; a
  // member's register sets start bstride floats after the previous
  // member's, its surface copies 6*ps after, its PML auxiliaries 24*cnt
  // after (per face)
  long bstride;
  int slot0, nm;
  int xg;  // items per XCD run (half the wide tiles of a row)
  // mode C (terrain-following grid, CURV instantiations): slopes p, q,
  // zeta_r, 1/r and the stage's one-sided zeta derivatives of p and q, per
  // cell in the field layout
  const float* __restrict__ cp;
  const float* __restrict__ cq;
  const float* __restrict__ czz;
  const float* __restrict__ crr;
  const float* __restrict__ cdp;
  const float* __restrict__ cdq;
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

// Auxiliary values of three variables of the first face slab holding a cell
// (see aux_plane), loaded with the plane's other loads: psi of the stage's
// input set, when FULL the set-A and set-T values the RK form reads, and the
// face axis' profile values 1/B, A + D/B, D/B.
struct AuxVals {
  float p[3], a[3], t[3];
  float b1, ad, dd;
  int face;   // 0..5, or -1: the cell is in no slab
  float* bp;     // set A, first of the three variables, this cell
  unsigned cnt;  // the face's cell count (24 cnt < 2^32: checked at set-up)
};
__device__ __forceinline__ float sel3(int m, float v0, float v1, float v2) { return m == 0 ? v0 : (m == 1 ? v1 : v2); }

// auxiliary RK update of variable m (0..2) of the three at bp (set A, this
// cell) of a face with cnt cells, in the RK form above (the reference's
// abs_RK_beg/inn/fin, mod_abs_npml.F90:738-1105, in 4-register form); psi is
// the stage input's value (W_1 / W_3 at forms 1 / 3).  pre: the set-A /
// set-T values were loaded ahead (pa, pt)
template <int FORM, bool FULL>
__device__ __forceinline__ void aux_rk_at(const StageArgs& g, float* bp, unsigned cnt, int m, float psi, float dpsi,
                                          bool pre, float pa, float pt) {
  // 32-bit offsets: a member's auxiliaries of a face span 24 cnt < 2^32 floats
  const unsigned S = 6u * cnt, q = (unsigned)m * cnt, so = (unsigned)g.setout * S;
  if constexpr (FORM == 0) {
    bp[so + q] = psi + g.a * dpsi;
  } else {
    if constexpr (FORM == 1) {
      const float A = FULL && pre ? pa : bp[q];
      bp[so + q] = A + g.a * dpsi;
      bp[3u * S + q] = g.c * A + g.r * psi + g.b * dpsi;
    } else if constexpr (FORM == 2) {
      const float A = FULL && pre ? pa : bp[q];
      bp[so + q] = A + g.a * dpsi;
    } else {
      const float T = FULL && pre ? pt : bp[3u * S + q];
      bp[so + q] = T + g.r * psi + g.b * dpsi;
    }
  }
}

// Prefetch of the auxiliary values of the first face slab holding a cell,
// in the reference's face order x1 x2 y1 y2 z1 z2.  x and y slabs span
// every plane of the volume, so their face and index step per plane are
// fixed for the column (AuxPre); a cell in no x or y slab takes the z slab
// of its plane.  The loads of plane k are issued with the RK operands,
// before the plane's barrier; cells in more than one slab (edges, corners)
// load the other slabs' values where they are applied.
struct AuxPre {
  float* base;        // set A, first of the three variables, this cell, plane k0
  int face;           // 0..3, or -1: no x or y slab
  int cnt, kst;       // face cell count, index step per plane
  int mask;           // bit n: the cell is in x or y slab n (0..3); bits 8, 9:
                      // the column is in the (i, j) extent of z slab 4, 5
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
  for (int n = 4; n < 6; ++n) {
    const FaceArgs& f = g.face[n];
    if (comp && f.on && i >= f.lo0 && i <= f.hi0 && j >= f.lo1 && j <= f.hi1) r.mask |= 1 << (n + 4);
  }
  return r;
}
// plane k lies in z slab n (4, 5): uniform over the block
__device__ __forceinline__ bool zplane(const StageArgs& g, int n, int k) {
  return g.face[n].on && k >= g.face[n].lo2 && k <= g.face[n].hi2;
}
// slabs holding cell (i, j, k) of a computed row: the column's x / y slabs
// and the plane's z slabs
__device__ __forceinline__ int face_mask(const StageArgs& g, const AuxPre& a, int i, int j, int k) {
  return (a.mask & 15) | ((a.mask & 256) && zplane(g, 4, k) ? 16 : 0) | ((a.mask & 512) && zplane(g, 5, k) ? 32 : 0);
}
template <int FORM, bool FULL>
__device__ __forceinline__ void aux_plane(const StageArgs& g, const AuxPre& a, int slot, int i, int j, int k, int k0, int v0,
                                          bool comp, AuxVals& r) {
  float* b = nullptr;
  unsigned c = 0;
  int ia = 0;
  r = AuxVals{{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, 0.f, 0.f, 0.f, -1, nullptr, 0u};
  if (a.face >= 0) {
    r.face = a.face;
    b = a.base + (unsigned)((k - k0) * a.kst);
    c = a.cnt;
    ia = a.face < 2 ? i : g.nx + j;
  } else if ((a.mask & 256) && zplane(g, 4, k)) {
    r.face = 4;
    c = g.face[4].cnt;
    b = aux_base(g.face[4], slot) + (long)v0 * c + face_index(g.face[4], i, j, k);
    ia = g.nx + g.ny + k;
  } else if ((a.mask & 512) && zplane(g, 5, k)) {
    r.face = 5;
    c = g.face[5].cnt;
    b = aux_base(g.face[5], slot) + (long)v0 * c + face_index(g.face[5], i, j, k);
    ia = g.nx + g.ny + k;
  }
  if (r.face >= 0) {
    r.bp = b;
    r.cnt = c;
    const unsigned in = (unsigned)g.setin * 6u * c;
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      r.p[m] = b[in + m * c];
      r.a[m] = FULL && (FORM == 1 || FORM == 2) ? b[m * c] : 0.f;
      r.t[m] = FULL && FORM == 3 ? b[18u * c + m * c] : 0.f;
    }
    r.b1 = g.pb1[ia];
    r.ad = g.pad[ia];
    r.dd = g.pdd[ia];
  }
}
// psi, set A and set T ahead in the wide tiles of both groups; narrow tiles:
// psi only
constexpr bool AUX_FULL_VEL = true, AUX_FULL_STR = true;

// One face slab's corrections of hV (ax: its axis) and the RK update of its
// three auxiliaries at bp (variables 3..5 of the face; pre: their set-A /
// set-T values were loaded ahead, in pv)
template <int FORM, bool FULL>
__device__ __forceinline__ void vel_face(const StageArgs& g, int ax, float b1, float ad, float Dd, float p3, float p4,
                                         float p5, float* bp, unsigned cnt, bool pre, const AuxVals& pv, float h[3], float rr,
                                         float DxTxx, float DxTxy, float DxTxz, float DyTyy, float DyTxy, float DyTyz,
                                         float DzTzz, float DzTxz, float DzTyz, float Txx, float Tyy, float Txy, float Tzz,
                                         float Txz, float Tyz, float rzk, float rxs, float xc) {
  const float b1_1 = b1 - 1.0f;
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
  aux_rk_at<FORM, FULL>(g, bp, cnt, v0 - 3, q0, Dd * D0 - ad * q0, pre, sel3(v0 - 3, pv.a[0], pv.a[1], pv.a[2]),
                        sel3(v0 - 3, pv.t[0], pv.t[1], pv.t[2]));
  aux_rk_at<FORM, FULL>(g, bp, cnt, v1 - 3, q1, Dd * D1 - ad * q1, pre, sel3(v1 - 3, pv.a[0], pv.a[1], pv.a[2]),
                        sel3(v1 - 3, pv.t[0], pv.t[1], pv.t[2]));
  aux_rk_at<FORM, FULL>(g, bp, cnt, v2 - 3, q2, Dd * D2 - ad * q2, pre, sel3(v2 - 3, pv.a[0], pv.a[1], pv.a[2]),
                        sel3(v2 - 3, pv.t[0], pv.t[1], pv.t[2]));
}

// CFS-PML corrections of hV on the face slabs containing (i, j, k), in the
// reference's face order x1 x2 y1 y2 z1 z2 (mod_abs_npml.F90:1149-2123, the
// stress-derivative half).  D* are the raw one-sided derivatives times the
// 1-D metric (before fluid zeroing).  The first slab holding the cell (pf)
// takes the values loaded with the plane (pv); the faces after it (cells at
// slab edges and corners only) load theirs here.
template <int FORM, bool FULL>
__device__ __forceinline__ void pml_vel(const StageArgs& g, int slot, int i, int j, int k, float h[3], float rr, float DxTxx,
                                        float DxTxy, float DxTxz, float DyTyy, float DyTxy, float DyTyz, float DzTzz,
                                        float DzTxz, float DzTyz, float Txx, float Tyy, float Txy, float Tzz, float Txz,
                                        float Tyz, float rzk, float rxs, float xc, int mask, int pf, const AuxVals pv) {
  if (pf >= 0)
    vel_face<FORM, FULL>(g, pf >> 1, pv.b1, pv.ad, pv.dd, pv.p[0], pv.p[1], pv.p[2], pv.bp, pv.cnt, true, pv, h, rr,
                         DxTxx, DxTxy, DxTxz, DyTyy, DyTxy, DyTyz, DzTzz, DzTxz, DzTyz, Txx, Tyy, Txy, Tzz, Txz, Tyz, rzk,
                         rxs, xc);
  const int rest = pf >= 0 ? mask & ~(1 << pf) : mask;
  if (rest == 0) return;
#pragma nounroll
  for (int n = 0; n < 6; ++n) {
    if (!((rest >> n) & 1)) continue;
    const FaceArgs& f = g.face[n];
    const int ax = n >> 1;
    const int ia = ax == 0 ? i : (ax == 1 ? g.nx + j : g.nx + g.ny + k);
    const unsigned cnt = (unsigned)f.cnt;
    float* const bp = aux_base(f, slot) + 3u * cnt + face_index(f, i, j, k);
    const unsigned in = (unsigned)g.setin * 6u * cnt;
    vel_face<FORM, FULL>(g, ax, g.pb1[ia], g.pad[ia], g.pdd[ia], bp[in], bp[in + cnt], bp[in + 2 * cnt], bp, cnt, false,
                         pv, h, rr, DxTxx, DxTxy, DxTxz, DyTyy, DyTxy, DyTyz, DzTzz, DzTxz, DzTyz, Txx, Tyy, Txy, Tzz, Txz,
                         Tyz, rzk, rxs, xc);
  }
}

// One face slab's corrections of hT (ax: its axis; surf: the free-surface
// Hooke fix-ups of the x and y faces, mod_abs_npml.F90:1303-1333) and the RK
// update of its three auxiliaries at bp (variables 0..2 of the face; pre:
// their set-A / set-T values were loaded ahead, in pv).
// h: Txx Tyy Tzz Txy Txz Tyz.  DzV* are the standard 5-tap z derivatives.
template <int FORM, bool FULL>
__device__ __forceinline__ void str_face(const StageArgs& g, int ax, bool surf, float b1, float ad, float Dd, float p0,
                                         float p1, float p2, float* bp, unsigned cnt, bool pre, const AuxVals& pv, float h[6],
                                         float lam, float miu, float DxVx, float DxVy, float DxVz, float DyVx, float DyVy,
                                         float DyVz, float DzVx, float DzVy, float DzVz, float Vx, float Vy, float Vz,
                                         float rzk, float rxs, float xc) {
  const float tc13 = lam, tc66 = miu, tc11 = tc13 + 2.0f * tc66;
  const float tc12 = tc13, tc22 = tc11, tc23 = tc13, tc33 = tc11, tc44 = tc66, tc55 = tc66;
  const float b1_1 = b1 - 1.0f;
  float D0, D1, D2;
  if (ax == 0) {
    D0 = (DxVx + Vz) * rzk; D1 = DxVy * rzk; D2 = (DxVz - Vx) * rzk;
    h[0] += b1_1 * tc11 * D0 - b1 * tc11 * p0;
    h[1] += b1_1 * tc12 * D0 - b1 * tc12 * p0;
    h[2] += b1_1 * tc13 * D0 - b1 * tc13 * p0;
    h[3] += b1_1 * tc66 * D1 - b1 * tc66 * p1;
    h[4] += b1_1 * tc55 * D2 - b1 * tc55 * p2;
    if (surf) {
      h[0] += -b1_1 * tc13 * tc13 / tc33 * D0 + b1 * tc13 * tc13 / tc33 * p0;
      h[1] += -b1_1 * tc23 * tc13 / tc33 * D0 + b1 * tc23 * tc13 / tc33 * p0;
    }
  } else if (ax == 1) {
    D0 = (DyVx * rxs - Vy * xc) * rzk;
    D1 = (DyVy * rxs + Vx * xc + Vz) * rzk;
    D2 = (DyVz * rxs - Vy) * rzk;
    h[3] += b1_1 * tc66 * D0 - b1 * tc66 * p0;
    h[0] += b1_1 * tc12 * D1 - b1 * tc12 * p1;
    h[1] += b1_1 * tc22 * D1 - b1 * tc22 * p1;
    h[2] += b1_1 * tc23 * D1 - b1 * tc23 * p1;
    h[5] += b1_1 * tc44 * D2 - b1 * tc44 * p2;
    if (surf) {
      h[0] += -b1_1 * tc13 * tc23 / tc33 * D1 + b1 * tc13 * tc23 / tc33 * p1;
      h[1] += -b1_1 * tc23 * tc23 / tc33 * D1 + b1 * tc23 * tc23 / tc33 * p1;
    }
  } else {
    D0 = DzVx; D1 = DzVy; D2 = DzVz;
    h[4] += b1_1 * tc55 * DzVx - b1 * tc55 * p0;
    h[5] += b1_1 * tc44 * DzVy - b1 * tc44 * p1;
    h[0] += b1_1 * tc13 * DzVz - b1 * tc13 * p2;
    h[1] += b1_1 * tc23 * DzVz - b1 * tc23 * p2;
    h[2] += b1_1 * tc33 * DzVz - b1 * tc33 * p2;
  }
  aux_rk_at<FORM, FULL>(g, bp, cnt, 0, p0, Dd * D0 - ad * p0, pre, pv.a[0], pv.t[0]);
  aux_rk_at<FORM, FULL>(g, bp, cnt, 1, p1, Dd * D1 - ad * p1, pre, pv.a[1], pv.t[1]);
  aux_rk_at<FORM, FULL>(g, bp, cnt, 2, p2, Dd * D2 - ad * p2, pre, pv.a[2], pv.t[2]);
}

// CFS-PML corrections of hT on the face slabs containing (i, j, k), in the
// reference's face order; the first slab holding the cell (pf) takes the
// values loaded with the plane (pv), the faces after it (slab edges and
// corners only) load theirs here.
template <int FORM, bool FULL>
__device__ __forceinline__ void pml_str(const StageArgs& g, int slot, int i, int j, int k, float h[6], float lam, float miu,
                                        float DxVx, float DxVy, float DxVz, float DyVx, float DyVy, float DyVz,
                                        float DzVx, float DzVy, float DzVz, float Vx, float Vy, float Vz, float rzk,
                                        float rxs, float xc, int mask, int pf, const AuxVals pv) {
  const bool surf = g.freenode && k == g.nk2;
  if (pf >= 0)
    str_face<FORM, FULL>(g, pf >> 1, surf, pv.b1, pv.ad, pv.dd, pv.p[0], pv.p[1], pv.p[2], pv.bp, pv.cnt, true, pv, h,
                         lam, miu, DxVx, DxVy, DxVz, DyVx, DyVy, DyVz, DzVx, DzVy, DzVz, Vx, Vy, Vz, rzk, rxs, xc);
  const int rest = pf >= 0 ? mask & ~(1 << pf) : mask;
  if (rest == 0) return;
#pragma nounroll
  for (int n = 0; n < 6; ++n) {
    if (!((rest >> n) & 1)) continue;
    const FaceArgs& f = g.face[n];
    const int ax = n >> 1;
    const int ia = ax == 0 ? i : (ax == 1 ? g.nx + j : g.nx + g.ny + k);
    const unsigned cnt = (unsigned)f.cnt;
    float* const bp = aux_base(f, slot) + face_index(f, i, j, k);
    const unsigned in = (unsigned)g.setin * 6u * cnt;
    str_face<FORM, FULL>(g, ax, surf, g.pb1[ia], g.pad[ia], g.pdd[ia], bp[in], bp[in + cnt], bp[in + 2 * cnt], bp, cnt,
                         false, pv, h, lam, miu, DxVx, DxVy, DxVz, DyVx, DyVy, DyVz, DzVx, DzVy, DzVz, Vx, Vy, Vz, rzk,
                         rxs, xc);
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

// Halo slot of one thread of a group in one LDS plane: x halo (XL left + XR
// right columns of the TY rows, threads tid < 4*TY) or y halo (YL + YR rows
// of the TW columns, every thread).  Slots outside the volume or beyond what
// the tile's computed columns ie / rows je read (partial tiles) hold zero
// and never feed a computed cell.
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
// Loads without branches: every lane's address is valid (lanes outside the
// tile or the volume read column 0 of their plane, halo slots outside read
// offset 0), and the value is selected afterwards.  The loads of a plane then
// form one basic block that the compiler schedules as a whole, instead of one
// exec-masked block per load.
__device__ __forceinline__ float ldv(const float* __restrict__ p, unsigned o, bool ok) {
  const float v = p[o];
  return ok ? v : 0.f;
}
__device__ __forceinline__ float ld(const float* __restrict__ p, const HaloSlot& h, unsigned kz) {
  return ldv(p, h.off + kz, h.ok);
}

// Mode C: the traction combinations across the surface zeta = const of a
// cell at offset a on plane kk, tau = sigma . N with N = (p, q, 1):
//   tau_z = Tzz + p Txz + q Tyz, tau_x = Txz + p Txx + q Txy, tau_y = Tyz + p Txy + q Tyy;
// above a free surface the imaged planes hold tau itself in Tzz Txz Tyz.
struct Sig6 {
  float s[6];  // Txx Tyy Txy Tzz Txz Tyz of one cell
  float p, q;
};
__device__ __forceinline__ Sig6 sig_at(const StageArgs& g, const float* __restrict__ pTxx, const float* __restrict__ pTyy,
                                       const float* __restrict__ pTxy, const float* __restrict__ pTzz,
                                       const float* __restrict__ pTxz, const float* __restrict__ pTyz, unsigned a,
                                       bool ok) {
  Sig6 r;
  r.s[0] = ldv(pTxx, a, ok); r.s[1] = ldv(pTyy, a, ok); r.s[2] = ldv(pTxy, a, ok);
  r.s[3] = ldv(pTzz, a, ok); r.s[4] = ldv(pTxz, a, ok); r.s[5] = ldv(pTyz, a, ok);
  r.p = ldv(g.cp, a, ok);
  r.q = ldv(g.cq, a, ok);
  return r;
}
__device__ __forceinline__ void tau_of(const StageArgs& g, const Sig6& v, int kk, float& tz, float& tx, float& ty) {
  const float xx = v.s[0], yy = v.s[1], xy = v.s[2], zz = v.s[3], xz = v.s[4], yz = v.s[5], p = v.p, q = v.q;
  if (g.freenode && kk > g.nk2) {
    tz = zz; tx = xz; ty = yz;
  } else {
    tz = zz + p * xz + q * yz;
    tx = xz + p * xx + q * xy;
    ty = yz + p * xy + q * yy;
  }
}
// Mode C FIFOs (one row per thread: the LDS budget).
// Stresses: a velocity-group thread keeps the six stresses of its cell from
// the window's lead plane until that plane becomes the centre plane (3
// planes later for F, 1 for B) in LDS words of its own, instead of loading
// them again (by then the lines have often left L2).  Layout [plane mod
// FD][field][thread]: consecutive threads hit consecutive banks.  FD = 3:
// at the end of iteration k the thread reads plane k+1 before it writes
// plane k+4 (F) into the same slot.
// Slopes: the velocity-group thread also writes p and q of the lead plane:
// the five planes of its z window stay there (FS = 5), from which it forms
// F(p) and F(q) itself, and the stress-group thread of the same cell (same
// thread index in its group) reads plane k after iteration k's barrier.  A
// slot is rewritten (plane k+5 at the end of iteration k+1 for F) only
// after the barrier of iteration k+1 that follows that read.
struct Fifo {
  static constexpr int FD = 3, FS = 5;
};

// LDS planes of the two groups (per buffer): velocity group Txx Txy Txz Tyy
// Tyz (with halos) and Tzz (centres only); stress group Vx Vy Vz.
template <int BYT, int RY, int TW>
struct TileGeom {
  static constexpr int TY = BYT * (TX / TW) * RY, LW = TW + 6, LH = TY + 6, PL = LW * LH;
};
constexpr int VNL = 6, SNL = 3;
enum { LTXX = 0, LTXY, LTXZ, LTYY, LTYZ, LTZZ };

// ---------------------------------------------------------------------------
// Velocity group: momentum equations with the spherical geometric terms
// (mod_macdrp.F90:928-933), fluid zeroing of shear-stress derivatives
// (:917-926), RK update of Vx Vy Vz.  z window: Tzz Txz Tyz.
// own: this group's LDS (VNL planes x 2 buffers), other: the stress group's.
// CURV (mode C, terrain-following grid): the z window holds the traction
// combinations tau_z tau_x tau_y (tau_at), the zeta part of row i of the
// stress divergence is zeta_r (F(tau_i) - sigma_ix F(p) - sigma_iy F(q)),
// 1/r is per cell; the plane centres of Tzz Txz Tyz are loaded separately.
// ---------------------------------------------------------------------------
template <int XF, int YF, int ZF, int FORM, int BYT, int RY, int TW, bool PML, bool CURV>
__device__ __forceinline__ void vel_body(const StageArgs& g, const int4 it, int slot, float* own, const float* other, int wy,
                                         float* fifo) {
  // RS: rows per row group (the stride between a thread's rows)
  using G = TileGeom<BYT, RY, TW>;
  constexpr int RS = BYT * (TX / TW), TY = G::TY, LW = G::LW, PL = G::PL;
  // set-A / set-T values of the CFS-PML auxiliaries loaded ahead in wide tiles
  constexpr bool FULLV = AUX_FULL_VEL && TW == TX;
  const int tx = threadIdx.x % TW, ty = threadIdx.x / TW + (TX / TW) * wy, tid = threadIdx.x + TX * wy;
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
  // form 3 writes set A in place: each cell's m is read, then overwritten,
  // by the same thread through the one pointer
  const float* __restrict__ mm = FORM == 3 ? wout : g.m + bo;
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

  float w[RY][3][5];  // Tzz Txz Tyz (CURV: tau_z tau_x tau_y)
  float cTxx[RY], cTyy[RY], cTxy[RY];
  float cTzz[RY], cTxz[RY], cTyz[RY];  // CURV: plane centres of the window fields
  // CURV with one row per thread: stresses from the lead plane to the centre
  // plane through the thread's own LDS words (FIFO)
  constexpr bool FIFO = CURV && RY == 1;
  constexpr int FD = Fifo::FD, FS = Fifo::FS, GT = TX * BYT;
  const int ftid = threadIdx.x + TX * wy;
  auto fword = [&](int plane, int f) -> float* { return fifo + ((plane % FD) * 6 + f) * GT + ftid; };
  // the slopes' FIFO for the stress group follows the stress FIFO
  auto sword = [&](int plane, int f) -> float* { return fifo + (FD * 6 + (plane % FS) * 2 + f) * GT + ftid; };
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
      if constexpr (CURV) {
        const Sig6 v = sig_at(g, pTxx, pTyy, pTxy, pTzz, pTxz, pTyz, c0 + o * (int)sz, ok[q]);
        tau_of(g, v, k0 + o, w[q][0][s], w[q][1][s], w[q][2][s]);
        if constexpr (FIFO) {
          if (o == 0) {
            cTxx[q] = v.s[0]; cTyy[q] = v.s[1]; cTxy[q] = v.s[2];
            cTzz[q] = v.s[3]; cTxz[q] = v.s[4]; cTyz[q] = v.s[5];
          } else if (o > 0) {
#pragma unroll
            for (int f = 0; f < 6; ++f) *fword(k0 + o, f) = v.s[f];
          }
          *sword(k0 + o, 0) = v.p;
          *sword(k0 + o, 1) = v.q;
        }
      } else {
        w[q][0][s] = ldv(pTzz, c0 + o * (int)sz, ok[q]);
        w[q][1][s] = ldv(pTxz, c0 + o * (int)sz, ok[q]);
        w[q][2][s] = ldv(pTyz, c0 + o * (int)sz, ok[q]);
      }
    }
    if constexpr (!FIFO) {
      cTxx[q] = ldv(pTxx, c0, ok[q]);
      cTyy[q] = ldv(pTyy, c0, ok[q]);
      cTxy[q] = ldv(pTxy, c0, ok[q]);
    }
    if constexpr (CURV && !FIFO) {
      cTzz[q] = ldv(pTzz, c0, ok[q]);
      cTxz[q] = ldv(pTxz, c0, ok[q]);
      cTyz[q] = ldv(pTyz, c0, ok[q]);
    }
  }
  AuxPre pre[RY];
#pragma unroll
  for (int q = 0; q < RY; ++q) pre[q] = PML ? aux_pre(g, slot, i, j0 + ty + RS * q, k0, 3, comp[q]) : AuxPre{nullptr, -1, 0, 0};
  unsigned kz = sz * (unsigned)k0;
  float hxTxx = ld(pTxx, hx, kz), hxTxy = ld(pTxy, hx, kz), hxTxz = ld(pTxz, hx, kz);
  float hyTxy = ld(pTxy, hy, kz), hyTyy = ld(pTyy, hy, kz), hyTyz = ld(pTyz, hy, kz);
  // halo lanes without a slot write a private scratch word of their own
  float* const spare = own + 2 * VNL * PL + (threadIdx.x & (TX - 1));
  float nrr[RY], nmo[RY][3], nto[RY][3];
  auto load_rk = [&](unsigned kp) {
#pragma unroll
    for (int q = 0; q < RY; ++q) {
      const unsigned c = cc[q] + kp;
      nrr[q] = lds_nt(g.rrho + c);
#pragma unroll
      for (int f = 0; f < 3; ++f) {
        const long e = f * fs + c;
        nmo[q][f] = FORM == 1 || FORM == 2 ? lds_nt(mm + e) : 0.f;
        nto[q][f] = FORM == 3 ? lds_nt(tt + e) : 0.f;
      }
    }
  };
  if (RK_AHEAD) load_rk(kz);

  for (int k = k0; k <= k1; ++k, kz += sz) {
    float* s = own + ((k - k0) & 1) * (VNL * PL);
    const float* sv = other + ((k - k0) & 1) * (SNL * PL);  // Vx Vy Vz of the stage input
    float* sTxx = s + LTXX * PL;
    float* sTxy = s + LTXY * PL;
    float* sTxz = s + LTXZ * PL;
    float* sTyy = s + LTYY * PL;
    float* sTyz = s + LTYZ * PL;
    float* sTzz = s + LTZZ * PL;
#pragma unroll
    for (int q = 0; q < RY; ++q) {
      const int o = (ty + RS * q + 3) * LW + tx + 3;
      sTxx[o] = cTxx[q];
      sTxy[o] = cTxy[q];
      sTyy[o] = cTyy[q];
      sTxz[o] = CURV ? cTxz[q] : zc<ZF>(w[q][1]);
      sTyz[o] = CURV ? cTyz[q] : zc<ZF>(w[q][2]);
      sTzz[o] = CURV ? cTzz[q] : zc<ZF>(w[q][0]);
    }
    {
      float* const px = hx.lds >= 0 ? s + hx.lds : spare;
      const int dx = hx.lds >= 0 ? PL : 0;
      px[LTXX * dx] = hxTxx;
      px[LTXY * dx] = hxTxy;
      px[LTXZ * dx] = hxTxz;
      float* const py = hy.lds >= 0 ? s + hy.lds : spare;
      const int dy = hy.lds >= 0 ? PL : 0;
      py[LTXY * dy] = hyTxy;
      py[LTYY * dy] = hyTyy;
      py[LTYZ * dy] = hyTyz;
    }
    // RK operands of plane k (form 0: m is the stage input, read from the
    // stress group's plane after the barrier); with RK_AHEAD those of plane
    // k+1 are loaded here and plane k's come from the previous iteration
    const bool more = k < k1;
    const unsigned kn = more ? kz + sz : kz;
    float rr[RY], mo[RY][3], to[RY][3];
#pragma unroll
    for (int q = 0; q < RY; ++q) {
      rr[q] = nrr[q];
#pragma unroll
      for (int f = 0; f < 3; ++f) { mo[q][f] = nmo[q][f]; to[q][f] = nto[q][f]; }
    }
    if (RK_AHEAD) load_rk(kn); else load_rk(kz);
    if (!RK_AHEAD) {
#pragma unroll
      for (int q = 0; q < RY; ++q) {
        rr[q] = nrr[q];
#pragma unroll
        for (int f = 0; f < 3; ++f) { mo[q][f] = nmo[q][f]; to[q][f] = nto[q][f]; }
      }
    }
    // prefetch plane k+1 (the last plane re-reads plane k: every address valid)
    float nTxx[RY], nTyy[RY], nTxy[RY], lead[RY][3];
    float nTzz[RY], nTxz[RY], nTyz[RY];
    Sig6 lsig[RY];  // CURV: the lead plane's stresses and slopes
    const int kl = (more ? k + 1 : k) + (ZF ? 3 : 1);
#pragma unroll
    for (int q = 0; q < RY; ++q) {
      const unsigned c = cc[q] + kn;
      const unsigned l = c + (ZF ? 3u : 1u) * sz;
      if constexpr (CURV) {
        lsig[q] = sig_at(g, pTxx, pTyy, pTxy, pTzz, pTxz, pTyz, l, ok[q]);
        if constexpr (!FIFO) {
          nTzz[q] = ldv(pTzz, c, ok[q]);
          nTxz[q] = ldv(pTxz, c, ok[q]);
          nTyz[q] = ldv(pTyz, c, ok[q]);
        }
      } else {
        lead[q][0] = ldv(pTzz, l, ok[q]);
        lead[q][1] = ldv(pTxz, l, ok[q]);
        lead[q][2] = ldv(pTyz, l, ok[q]);
      }
      if constexpr (!FIFO) {
        nTxx[q] = ldv(pTxx, c, ok[q]);
        nTyy[q] = ldv(pTyy, c, ok[q]);
        nTxy[q] = ldv(pTxy, c, ok[q]);
      }
    }
    const float nhxTxx = ld(pTxx, hx, kn), nhxTxy = ld(pTxy, hx, kn), nhxTxz = ld(pTxz, hx, kn);
    const float nhyTxy = ld(pTxy, hy, kn), nhyTyy = ld(pTyy, hy, kn), nhyTyz = ld(pTyz, hy, kn);
    AuxVals pv[RY];
    if constexpr (PML) {
#pragma unroll
      for (int q = 0; q < RY; ++q) aux_plane<FORM, FULLV>(g, pre[q], slot, i, j0 + ty + RS * q, k, k0, 3, comp[q], pv[q]);
    }
    // plane metric with the plane's other loads (used right after the barrier)
    const float ze1 = g.zeta_z[k], rz1 = g.rz[k];
    float mze[RY], mrz[RY], mdp[RY], mdq[RY];  // CURV: zeta_r, 1/r, F(p), F(q) of the cell
    if constexpr (CURV) {
#pragma unroll
      for (int q = 0; q < RY; ++q) {
        const unsigned c = cc[q] + kz;
        mze[q] = ldv(g.czz, c, ok[q]);
        mrz[q] = ldv(g.crr, c, ok[q]);
        if constexpr (!FIFO) {
          mdp[q] = ldv(g.cdp, c, ok[q]);
          mdq[q] = ldv(g.cdq, c, ok[q]);
        }
      }
    }
    __syncthreads();
    const bool pml_here = PML && (tile_pml || (g.face[4].on && k <= g.face[4].hi2) || (g.face[5].on && k >= g.face[5].lo2));
#pragma unroll
    for (int q = 0; q < RY; ++q) {
      if (comp[q]) {
        const unsigned c = cc[q] + kz;
        const int o = (ty + RS * q + 3) * LW + tx + 3;
        const float ze = CURV ? mze[q] : ze1, rzk = CURV ? mrz[q] : rz1;
        const float DxTxx = dlds<XF>(sTxx + o, 1) * xi;
        float DxTxy = dlds<XF>(sTxy + o, 1) * xi;
        float DxTxz = dlds<XF>(sTxz + o, 1) * xi;
        const float DyTyy = dlds<YF>(sTyy + o, LW) * et[q];
        float DyTxy = dlds<YF>(sTxy + o, LW) * et[q];
        float DyTyz = dlds<YF>(sTyz + o, LW) * et[q];
        const float Txx = cTxx[q], Tyy = cTyy[q], Txy = cTxy[q];
        const float Tzz = CURV ? cTzz[q] : zc<ZF>(w[q][0]), Txz = CURV ? cTxz[q] : zc<ZF>(w[q][1]),
                    Tyz = CURV ? cTyz[q] : zc<ZF>(w[q][2]);
        float DzTzz0, DzTxz0, DzTyz0;
        if constexpr (FIFO) {
          // F(p), F(q) from the slopes of the window's planes
          float pw[5], qw[5];
#pragma unroll
          for (int sl = 0; sl < 5; ++sl) {
            pw[sl] = *sword(k + (ZF ? sl - 1 : sl - 3), 0);
            qw[sl] = *sword(k + (ZF ? sl - 1 : sl - 3), 1);
          }
          mdp[q] = dz5<ZF>(pw);
          mdq[q] = dz5<ZF>(qw);
        }
        if constexpr (CURV) {
          DzTzz0 = (dz5<ZF>(w[q][0]) - Txz * mdp[q] - Tyz * mdq[q]) * ze;
          DzTxz0 = (dz5<ZF>(w[q][1]) - Txx * mdp[q] - Txy * mdq[q]) * ze;
          DzTyz0 = (dz5<ZF>(w[q][2]) - Txy * mdp[q] - Tyy * mdq[q]) * ze;
        } else {
          DzTzz0 = dz5<ZF>(w[q][0]) * ze;
          DzTxz0 = dz5<ZF>(w[q][1]) * ze;
          DzTyz0 = dz5<ZF>(w[q][2]) * ze;
        }
        const float DzTzz = DzTzz0;
        float DzTxz = DzTxz0, DzTyz = DzTyz0;
        const float rDxTxy = DxTxy, rDxTxz = DxTxz, rDyTxy = DyTxy, rDyTyz = DyTyz, rDzTxz = DzTxz, rDzTyz = DzTyz;
        // fluid cells (mu <= SEIS_ZERO) occur only in the rare columns that hold one
        if (fluid[q] && g.mu[c] <= SEIS_ZERO) {
          DxTxy = 0.f; DxTxz = 0.f; DyTxy = 0.f; DyTyz = 0.f; DzTxz = 0.f; DzTyz = 0.f;
        }
        float h[3];
        h[0] = rr[q] * (DxTxx * rzk + DyTxy * rzk * rxs + DzTxz + (3.0f * Txz + Txx * xc - Tyy * xc) * rzk);
        h[1] = rr[q] * (DxTxy * rzk + DyTyy * rzk * rxs + DzTyz + (2.0f * Txy * xc + 3.0f * Tyz) * rzk);
        h[2] = rr[q] * (DxTxz * rzk + DyTyz * rzk * rxs + DzTzz + (2.0f * Tzz - Txx - Tyy + Txz * xc) * rzk);
        if constexpr (PML)
          if (pml_here) pml_vel<FORM, FULLV>(g, slot, i, j0 + ty + RS * q, k, h, rr[q], DxTxx, rDxTxy, rDxTxz, DyTyy, rDyTxy, rDyTyz, DzTzz,
                        rDzTxz, rDzTyz, Txx, Tyy, Txy, Tzz, Txz, Tyz, rzk, rxs, xc, face_mask(g, pre[q], i, j0 + ty + RS * q, k), pv[q].face, pv[q]);
#pragma unroll
        for (int f = 0; f < 3; ++f) {
          const long e = f * fs + c;
          const float wc = sv[f * PL + o];  // the stage input's own value (velocities are never imaged)
          if constexpr (FORM == 0) {
            st_nt(wout + e, wc + g.a * h[f]);
          } else if constexpr (FORM == 1) {
            st_nt(wout + e, mo[q][f] + g.a * h[f]);
            st_nt(tt + e, g.c * mo[q][f] + g.r * wc + g.b * h[f]);
          } else if constexpr (FORM == 2) {
            st_nt(wout + e, mo[q][f] + g.a * h[f]);
          } else {
            st_nt(wout + e, to[q][f] + g.r * wc + g.b * h[f]);
          }
        }
      }
      // advance the window and the plane registers
      if constexpr (CURV) tau_of(g, lsig[q], kl, lead[q][0], lead[q][1], lead[q][2]);
#pragma unroll
      for (int f = 0; f < 3; ++f) {
#pragma unroll
        for (int sl = 0; sl < 4; ++sl) w[q][f][sl] = w[q][f][sl + 1];
        w[q][f][4] = lead[q][f];
      }
      if constexpr (FIFO) {
        // the next centre plane's stresses out, then the lead plane's in
        // (for F the same slot), and the lead plane's slopes for the stress
        // group
        cTxx[q] = *fword(k + 1, 0); cTyy[q] = *fword(k + 1, 1); cTxy[q] = *fword(k + 1, 2);
        cTzz[q] = *fword(k + 1, 3); cTxz[q] = *fword(k + 1, 4); cTyz[q] = *fword(k + 1, 5);
#pragma unroll
        for (int f = 0; f < 6; ++f) *fword(kl, f) = lsig[q].s[f];
        *sword(kl, 0) = lsig[q].p;
        *sword(kl, 1) = lsig[q].q;
      } else {
        cTxx[q] = nTxx[q];
        cTyy[q] = nTyy[q];
        cTxy[q] = nTxy[q];
        if constexpr (CURV) {
          cTzz[q] = nTzz[q];
          cTxz[q] = nTxz[q];
          cTyz[q] = nTyz[q];
        }
      }
    }
    hxTxx = nhxTxx; hxTxy = nhxTxy; hxTxz = nhxTxz;
    hyTxy = nhyTxy; hyTyy = nhyTyy; hyTyz = nhyTyz;
  }
}

// ---------------------------------------------------------------------------
// Stress group: strain rates (mod_macdrp.F90:935-941), near-surface low-order
// z operators (CondFreeVLOW, :872-915), the traction condition at k = nk2
// with zero source terms (:942-950; k_src adds the source part), Hooke
// (:952-957), RK update of the six stresses.  At k = nk2 the "m" values of
// Tzz Txz Tyz are the pre-image stresses saved by k_surface at form 0, and
// their W_1 / W_3 the pre-image inputs saved at forms 1 / 3.
// z window: Vx Vy Vz.  own: this group's LDS (SNL planes x 2 buffers),
// other: the velocity group's.
// CURV (mode C): with a = zeta_r F(V) the strain rates take the slope terms
//   E11 += p aX, E22 += q aY, E12 += (p aY + q aX)/2, E13 += p aZ/2, E23 += q aZ/2,
// and at a free-surface cell with p or q nonzero a comes from the
// traction-rate condition (surf_solve; the source part is added by k_src).
// ---------------------------------------------------------------------------
template <int XF, int YF, int ZF, int FORM, int BYT, int RY, int TW, bool PML, bool CURV>
__device__ __forceinline__ void str_body(const StageArgs& g, const int4 it, int slot, float* own, const float* other, int wy,
                                         const float* fifo) {
  using G = TileGeom<BYT, RY, TW>;
  constexpr int RS = BYT * (TX / TW), TY = G::TY, LW = G::LW, PL = G::PL;
  constexpr bool FULLS = AUX_FULL_STR && TW == TX;
  // LDS plane of each output field (Txx Tyy Tzz Txy Txz Tyz) in the velocity group's planes
  constexpr int LP[6] = {LTXX, LTYY, LTZZ, LTXY, LTXZ, LTYZ};
  const int tx = threadIdx.x % TW, ty = threadIdx.x / TW + (TX / TW) * wy, tid = threadIdx.x + TX * wy;
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
  const float* __restrict__ mm = FORM == 3 ? wout : g.m + bo;  // as in the velocity group
  float* __restrict__ tt = g.t + bo;
  const float* __restrict__ ms = g.msurf + (long)slot * 6 * g.ps;
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
      w[q][0][s] = ldv(pV0, c0 + o * (int)sz, ok[q]);
      w[q][1][s] = ldv(pV1, c0 + o * (int)sz, ok[q]);
      w[q][2][s] = ldv(pV2, c0 + o * (int)sz, ok[q]);
    }
  }
  AuxPre pre[RY];
#pragma unroll
  for (int q = 0; q < RY; ++q) pre[q] = PML ? aux_pre(g, slot, i, j0 + ty + RS * q, k0, 0, comp[q]) : AuxPre{nullptr, -1, 0, 0};
  unsigned kz = sz * (unsigned)k0;
  float hx0 = ld(pV0, hx, kz), hx1 = ld(pV1, hx, kz), hx2 = ld(pV2, hx, kz);
  float hy0 = ld(pV0, hy, kz), hy1 = ld(pV1, hy, kz), hy2 = ld(pV2, hy, kz);
  // halo lanes without a slot write a private scratch word of their own
  float* const spare = own + 2 * SNL * PL + (threadIdx.x & (TX - 1));
  float nla[RY], nmu[RY], nmo[RY][6], nto[RY][6];
  auto load_rk = [&](unsigned kp) {
#pragma unroll
    for (int q = 0; q < RY; ++q) {
      const unsigned c = cc[q] + kp;
      nla[q] = lds_nt(g.lam + c);
      nmu[q] = lds_nt(g.mu + c);
#pragma unroll
      for (int f = 0; f < 6; ++f) {
        const long e = (TXX + f) * fs + c;
        nmo[q][f] = FORM == 1 || FORM == 2 ? lds_nt(mm + e) : 0.f;
        nto[q][f] = FORM == 3 ? lds_nt(tt + e) : 0.f;
      }
    }
  };
  if (RK_AHEAD) load_rk(kz);

  for (int k = k0; k <= k1; ++k, kz += sz) {
    float* s = own + ((k - k0) & 1) * (SNL * PL);
    const float* st = other + ((k - k0) & 1) * (VNL * PL);  // stresses of the stage input
#pragma unroll
    for (int q = 0; q < RY; ++q) {
      const int o = (ty + RS * q + 3) * LW + tx + 3;
      s[o] = zc<ZF>(w[q][0]);
      s[PL + o] = zc<ZF>(w[q][1]);
      s[2 * PL + o] = zc<ZF>(w[q][2]);
    }
    {
      float* const px = hx.lds >= 0 ? s + hx.lds : spare;
      const int dx = hx.lds >= 0 ? PL : 0;
      px[0] = hx0;
      px[dx] = hx1;
      px[2 * dx] = hx2;
      float* const py = hy.lds >= 0 ? s + hy.lds : spare;
      const int dy = hy.lds >= 0 ? PL : 0;
      py[0] = hy0;
      py[dy] = hy1;
      py[2 * dy] = hy2;
    }
    const bool surf = g.freenode && k == g.nk2;
    // RK operands of plane k (form 0: m is the stage input, read from the
    // velocity group's plane after the barrier; the surface copies of the
    // imaged fields are read there too); with RK_AHEAD as in the velocity group
    const bool more = k < k1;
    const unsigned kn = more ? kz + sz : kz;
    float la[RY], mu_[RY], mo[RY][6], to[RY][6];
#pragma unroll
    for (int q = 0; q < RY; ++q) {
      la[q] = nla[q]; mu_[q] = nmu[q];
#pragma unroll
      for (int f = 0; f < 6; ++f) { mo[q][f] = nmo[q][f]; to[q][f] = nto[q][f]; }
    }
    if (RK_AHEAD) load_rk(kn); else load_rk(kz);
    if (!RK_AHEAD) {
#pragma unroll
      for (int q = 0; q < RY; ++q) {
        la[q] = nla[q]; mu_[q] = nmu[q];
#pragma unroll
        for (int f = 0; f < 6; ++f) { mo[q][f] = nmo[q][f]; to[q][f] = nto[q][f]; }
      }
    }
    // prefetch plane k+1 (the last plane re-reads plane k: every address valid)
    float lead[RY][3];
#pragma unroll
    for (int q = 0; q < RY; ++q) {
      const unsigned l = cc[q] + kn + (ZF ? 3u : 1u) * sz;
      lead[q][0] = ldv(pV0, l, ok[q]);
      lead[q][1] = ldv(pV1, l, ok[q]);
      lead[q][2] = ldv(pV2, l, ok[q]);
    }
    const float nx0 = ld(pV0, hx, kn), nx1 = ld(pV1, hx, kn), nx2 = ld(pV2, hx, kn);
    const float ny0 = ld(pV0, hy, kn), ny1 = ld(pV1, hy, kn), ny2 = ld(pV2, hy, kn);
    AuxVals pv[RY];
    if constexpr (PML) {
#pragma unroll
      for (int q = 0; q < RY; ++q) aux_plane<FORM, FULLS>(g, pre[q], slot, i, j0 + ty + RS * q, k, k0, 0, comp[q], pv[q]);
    }
    // plane metric with the plane's other loads (used right after the barrier)
    const float ze1 = g.zeta_z[k], rz1 = g.rz[k];
    float mze[RY], mrz[RY], mp[RY], mq[RY];  // CURV: zeta_r, 1/r, slopes of the cell
    // CURV with one row per thread: the slopes come from the velocity group's
    // FIFO (written at its window lead), read after the barrier
    constexpr bool SFIFO = CURV && RY == 1;
    if constexpr (CURV) {
#pragma unroll
      for (int q = 0; q < RY; ++q) {
        const unsigned c = cc[q] + kz;
        mze[q] = ldv(g.czz, c, ok[q]);
        mrz[q] = ldv(g.crr, c, ok[q]);
        if constexpr (!SFIFO) {
          mp[q] = ldv(g.cp, c, ok[q]);
          mq[q] = ldv(g.cq, c, ok[q]);
        }
      }
    }
    __syncthreads();
    if constexpr (SFIFO) {
      constexpr int FD = Fifo::FD, FS = Fifo::FS, GT = TX * BYT;
      const int ftid = threadIdx.x + TX * wy;
      mp[0] = fifo[(FD * 6 + (k % FS) * 2 + 0) * GT + ftid];
      mq[0] = fifo[(FD * 6 + (k % FS) * 2 + 1) * GT + ftid];
    }
    const bool pml_here = PML && (tile_pml || (g.face[4].on && k <= g.face[4].hi2) || (g.face[5].on && k >= g.face[5].lo2));
#pragma unroll
    for (int q = 0; q < RY; ++q) {
      if (comp[q]) {
        const unsigned c = cc[q] + kz;
        const int o = (ty + RS * q + 3) * LW + tx + 3;
        const float ze = CURV ? mze[q] : ze1, rzk = CURV ? mrz[q] : rz1;
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
        float E11, E22, E12, E33, E13, E23;
        if constexpr (CURV) {
          const float p = mp[q], qs = mq[q];
          if (surf && !(p == 0.f && qs == 0.f)) {
            // the stress rates without the zeta parts, then a from the
            // traction-rate condition with a zero source
            const float H11 = (DxVx + Vz) * rzk, H22 = (Vx * xc + DyVy * rxs + Vz) * rzk;
            const float H12 = (DxVy + DyVx * rxs - Vy * xc) * rzk * 0.5f;
            const float H13 = (DxVz * rzk - Vx * rzk) * 0.5f, H23 = (DyVz * rzk * rxs - Vy * rzk) * 0.5f;
            const float tr = H11 + H22;
            const float sxx = lam * tr + 2.0f * miu * H11, syy = lam * tr + 2.0f * miu * H22, szz = lam * tr;
            const float sxy = 2.0f * miu * H12, sxz = 2.0f * miu * H13, syz = 2.0f * miu * H23;
            float aX, aY, aZ;
            surf_solve(lam, miu, p, qs, -(p * sxx + qs * sxy + sxz), -(p * sxy + qs * syy + syz),
                       -(p * sxz + qs * syz + szz), aX, aY, aZ);
            E11 = H11 + p * aX;
            E22 = H22 + qs * aY;
            E33 = aZ;
            E12 = H12 + (p * aY + qs * aX) * 0.5f;
            E13 = H13 + (aX + p * aZ) * 0.5f;
            E23 = H23 + (aY + qs * aZ) * 0.5f;
          } else {
            E11 = (DxVx + Vz) * rzk + p * DzVx;
            E22 = (Vx * xc + DyVy * rxs + Vz) * rzk + qs * DzVy;
            E12 = ((DxVy + DyVx * rxs - Vy * xc) * rzk + p * DzVy + qs * DzVx) * 0.5f;
            E33 = DzVz;
            E13 = (DxVz * rzk + DzVx - Vx * rzk + p * DzVz) * 0.5f;
            E23 = (DyVz * rzk * rxs + DzVy - Vy * rzk + qs * DzVz) * 0.5f;
            if (surf) {
              E33 = -lam / lam2mu * (E11 + E22);
              E13 = 0.f;
              E23 = 0.f;
            }
          }
        } else {
          E11 = (DxVx + Vz) * rzk;
          E22 = (Vx * xc + DyVy * rxs + Vz) * rzk;
          E12 = (DxVy + DyVx * rxs - Vy * xc) * rzk * 0.5f;
          E33 = DzVz;
          E13 = (DxVz * rzk + DzVx - Vx * rzk) * 0.5f;
          E23 = (DyVz * rzk * rxs + DzVy - Vy * rzk) * 0.5f;
          if (surf) {
            E33 = -lam / lam2mu * (E11 + E22);
            E13 = 0.f;
            E23 = 0.f;
          }
        }
        float h[6];  // Txx Tyy Tzz Txy Txz Tyz
        h[0] = lam2mu * E11 + lam * (E22 + E33);
        h[1] = lam2mu * E22 + lam * (E11 + E33);
        h[2] = lam2mu * E33 + lam * (E11 + E22);
        h[3] = 2.0f * miu * E12;
        h[4] = 2.0f * miu * E13;
        h[5] = 2.0f * miu * E23;
        if constexpr (PML)
          if (pml_here) pml_str<FORM, FULLS>(g, slot, i, j0 + ty + RS * q, k, h, lam, miu, DxVx, DxVy, DxVz, DyVx, DyVy, DyVz, sDzVx, sDzVy,
                        sDzVz, Vx, Vy, Vz, rzk, rxs, xc, face_mask(g, pre[q], i, j0 + ty + RS * q, k), pv[q].face, pv[q]);
        if (surf) {
          // Tzz Txz Tyz are imaged at k = nk2: their m and W_1 / W_3 are the
          // copies k_surface saved before imaging
#pragma unroll
          for (int f = 2; f < 6; ++f) {
            if (f == 3) continue;
            const int sp = f == 2 ? 0 : (f == 4 ? 1 : 2);
            mo[q][f] = ms[sp * g.ps + cc[q]];
          }
        }
#pragma unroll
        for (int f = 0; f < 6; ++f) {
          const long e = (TXX + f) * fs + c;
          const bool img = surf && (f == 2 || f == 4 || f == 5);
          const int sp = f == 2 ? 0 : (f == 4 ? 1 : 2);
          float wc = st[LP[f] * PL + o];  // the stage input's own value
          // (the compiler merges the two reads into one flat load of a
          // selected address; keeping them apart measured 0.4% slower)
          if constexpr (FORM == 1 || FORM == 3)
            if (img) wc = ms[(3 + sp) * g.ps + cc[q]];
          if constexpr (FORM == 0) {
            st_nt(wout + e, (img ? mo[q][f] : wc) + g.a * h[f]);
          } else if constexpr (FORM == 1) {
            st_nt(wout + e, mo[q][f] + g.a * h[f]);
            st_nt(tt + e, g.c * mo[q][f] + g.r * wc + g.b * h[f]);
          } else if constexpr (FORM == 2) {
            st_nt(wout + e, mo[q][f] + g.a * h[f]);
          } else {
            st_nt(wout + e, to[q][f] + g.r * wc + g.b * h[f]);
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

// LDS of one block: the velocity group's planes, then the stress group's,
// each sized for the larger of the wide and narrow tiles.  Each group's region ends with TX spare words (halo writes of lanes
// without a halo slot).
template <int BYT, int RY, bool CURV>
struct StageLds {
  static constexpr int VW = 2 * VNL * TileGeom<BYT, RY, TX>::PL, VN = 2 * VNL * TileGeom<BYT, RY, TXN>::PL;
  static constexpr int SW = 2 * SNL * TileGeom<BYT, RY, TX>::PL, SN = 2 * SNL * TileGeom<BYT, RY, TXN>::PL;
  static constexpr int VOFF = (VW > VN ? VW : VN) + TX, FOFF = VOFF + (SW > SN ? SW : SN) + TX;
  // mode C FIFOs of the velocity group (one row per thread): stresses, then
  // the slopes handed to the stress group
  static constexpr int FF = (CURV && RY == 1) ? (Fifo::FD * 6 + Fifo::FS * 2) * TX * BYT : 0;
  static constexpr int SIZE = FOFF + FF;
};

// One launch per RK stage: blocks of 2 x BYT waves, the velocity group
// first.  The narrow edge tiles run the narrow body (with the CFS-PML
// corrections), the wide PML-shell tiles the PML body, the wide interior
// tiles the lean body.  At most 128 VGPRs (4 waves per SIMD: two 8-wave
// blocks per CU).
template <int XF, int YF, int ZF, int FORM, int BYT, int RY, bool CURV>
__global__ __launch_bounds__(2 * TX * BYT) __attribute__((amdgpu_waves_per_eu(4))) void k_stage(StageArgs g) {
  using L = StageLds<BYT, RY, CURV>;
  __shared__ float sh[L::SIZE];
  const ItemPos ip = item_pos(g);
  const int it_ = ip.item, slot = g.slot0 + ip.member;
  const int4 it = g.items[it_];
  // wave index, uniform over the wave: the group choice is a scalar branch
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.y);
  float* const sv = sh;
  float* const ss = sh + L::VOFF;
  if (wave < BYT) {
    if (it_ < g.p1)
      vel_body<XF, YF, ZF, FORM, BYT, RY, TXN, true, CURV>(g, it, slot, sv, ss, wave, sh + L::FOFF);
    else if (it_ < g.p2)
      vel_body<XF, YF, ZF, FORM, BYT, RY, TX, true, CURV>(g, it, slot, sv, ss, wave, sh + L::FOFF);
    else
      vel_body<XF, YF, ZF, FORM, BYT, RY, TX, false, CURV>(g, it, slot, sv, ss, wave, sh + L::FOFF);
  } else {
    if (it_ < g.p1)
      str_body<XF, YF, ZF, FORM, BYT, RY, TXN, true, CURV>(g, it, slot, ss, sv, wave - BYT, sh + L::FOFF);
    else if (it_ < g.p2)
      str_body<XF, YF, ZF, FORM, BYT, RY, TX, true, CURV>(g, it, slot, ss, sv, wave - BYT, sh + L::FOFF);
    else
      str_body<XF, YF, ZF, FORM, BYT, RY, TX, false, CURV>(g, it, slot, ss, sv, wave - BYT, sh + L::FOFF);
  }
}

// Stage-kernel tile shapes (the autotuner picks among them; the first is the
// default): waves per group x rows per thread = 4-row tiles (4 x 1), 8-row
// tiles (4 x 2, 8 x 1), 16-row tiles (8 x 2).
constexpr int NTILECFG = 4;
constexpr int TILE_BYT[NTILECFG] = {4, 4, 8, 8};  // waves per group (a block holds two groups)
constexpr int TILE_RY[NTILECFG] = {1, 2, 2, 1};   // rows per thread

// The 8 F/B patterns x 4 RK forms of one tile shape are instantiated in
// translation units of their own (fw_stage_inst.hip, compiled once per shape
// and pattern half, in parallel).
typedef void (*StageFn)(StageArgs);
StageFn stage_fn_0_0(int pat, int form);
StageFn stage_fn_0_1(int pat, int form);
StageFn stage_fn_1_0(int pat, int form);
StageFn stage_fn_1_1(int pat, int form);
StageFn stage_fn_2_0(int pat, int form);
StageFn stage_fn_2_1(int pat, int form);
StageFn stage_fn_3_0(int pat, int form);
StageFn stage_fn_3_1(int pat, int form);
// mode C (terrain-following grid) instantiations of the same shapes
StageFn stage_fnc_0_0(int pat, int form);
StageFn stage_fnc_0_1(int pat, int form);
StageFn stage_fnc_1_0(int pat, int form);
StageFn stage_fnc_1_1(int pat, int form);
StageFn stage_fnc_2_0(int pat, int form);
StageFn stage_fnc_2_1(int pat, int form);
StageFn stage_fnc_3_0(int pat, int form);
StageFn stage_fnc_3_1(int pat, int form);

}  // namespace fw
