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
// TX x TY tile of (i, j) columns over a chunk of k planes and marches k.
// The current k plane of every field that needs an x or y derivative is
// staged in LDS with its one-sided halo (F: -1..+3, B: -3..+1), double
// buffered so one barrier per plane suffices; z derivatives come from a
// 5-slot register window per column.  While plane k is computed, the loads
// of plane k+1 (window lead, halo) and the RK operands of plane k are in
// flight.  Wave w owns rows j0 + w + 4q (q < RY): every global access of a
// wave is 64 consecutive floats of one row.  Offsets are 32-bit relative to
// per-field base pointers.
#pragma once
#include "fw_common.h"

namespace fw {

constexpr int TX = 64;

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
  int kchunk, freenode;
  int ntx, nty, nch;
  float a, b;
};

// Blocks are dealt round-robin to the 8 XCDs; remap so that each XCD
// receives a contiguous run of logical tiles (shared halos stay in one L2).
__device__ __forceinline__ void tile_of_block(const StageArgs& g, int& bx, int& by, int& bz) {
  const int nb = g.ntx * g.nty * g.nch;
  const int b = blockIdx.x;
  const int q = nb / 8;
  const int L = (b < 8 * q) ? (b % 8) * q + b / 8 : b;
  bx = L % g.ntx;
  const int r = L / g.ntx;
  by = r % g.nty;
  bz = r / g.nty;
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
// TX columns, every thread).  Out-of-volume slots (partial edge tiles) hold
// zero and never feed a computed cell.
struct HaloSlot {
  unsigned off;
  int lds;
  bool ok;
};
template <int XF, int TY, int LW>
__device__ __forceinline__ HaloSlot xhalo_slot(const StageArgs& g, int i0, int j0, int tid) {
  constexpr int XL = XF ? 1 : 3;
  HaloSlot h{0u, 0, false};
  if (tid < 4 * TY) {
    const int r = tid % TY, s = tid / TY;
    const int col = s < XL ? i0 - XL + s : i0 + TX + (s - XL);
    const int row = j0 + r;
    h.ok = col < g.nx && row < g.ny;
    h.off = h.ok ? (unsigned)(col + g.sy * row) : 0u;
    h.lds = (r + 3) * LW + (col - i0 + 3);
  } else {
    h.lds = -1;
  }
  return h;
}
template <int YF, int TY, int LW>
__device__ __forceinline__ HaloSlot yhalo_slot(const StageArgs& g, int i0, int j0, int tid) {
  constexpr int YL = YF ? 1 : 3;
  HaloSlot h;
  const int c = tid % TX, s = tid / TX;
  const int row = s < YL ? j0 - YL + s : j0 + TY + (s - YL);
  const int col = i0 + c;
  h.ok = col < g.nx && row < g.ny;
  h.off = h.ok ? (unsigned)(col + g.sy * row) : 0u;
  h.lds = (row - j0 + 3) * LW + (c + 3);
  return h;
}
__device__ __forceinline__ float ld(const float* __restrict__ p, const HaloSlot& h, unsigned kz) {
  return h.ok ? p[h.off + kz] : 0.f;
}

constexpr int NTH = 256, BYT = 4;

// ---------------------------------------------------------------------------
// k_vel: momentum equations with the spherical geometric terms
// (mod_macdrp.F90:928-933), fluid zeroing of shear-stress derivatives
// (:917-926), RK update of Vx Vy Vz.
// LDS: Txx (x), Txy (x, y), Txz (x), Tyy (y), Tyz (y); z window: Tzz Txz Tyz.
// ---------------------------------------------------------------------------
template <int XF, int YF, int ZF, int KIND, int RY>
__global__ __launch_bounds__(NTH) void k_vel(StageArgs g) {
  constexpr int TY = BYT * RY, LW = TX + 6, LH = TY + 6, PL = LW * LH, NL = 5;
  __shared__ float sh[2][NL * PL];  // Txx Txy Txz Tyy Tyz, double buffered
  const int tx = threadIdx.x, ty = threadIdx.y, tid = tx + TX * ty;
  int bx, by, bz;
  tile_of_block(g, bx, by, bz);
  const int i0 = g.ni1 + bx * TX, j0 = g.nj1 + by * TY;
  const int k0 = g.nk1 + bz * g.kchunk, k1 = min(k0 + g.kchunk - 1, g.nk2);
  const long fs = g.fs;
  const unsigned sz = (unsigned)g.sz;
  const float* __restrict__ W = g.win;
  const float* __restrict__ pTxx = W + TXX * fs;
  const float* __restrict__ pTyy = W + TYY * fs;
  const float* __restrict__ pTxy = W + TXY * fs;
  const float* __restrict__ pTzz = W + TZZ * fs;
  const float* __restrict__ pTxz = W + TXZ * fs;
  const float* __restrict__ pTyz = W + TYZ * fs;
  const int i = i0 + tx;
  const bool iok = i < g.nx, icomp = i <= g.ni2;
  const int ic = min(i, g.nx - 1);
  const float xi = g.xi_x[ic], rxs = g.rxsin[ic], xc = g.xcot[ic];
  const HaloSlot hx = xhalo_slot<XF, TY, LW>(g, i0, j0, tid);
  const HaloSlot hy = yhalo_slot<YF, TY, LW>(g, i0, j0, tid);

  float w[RY][3][5];  // Tzz Txz Tyz
  float cTxx[RY], cTyy[RY], cTxy[RY];
  bool ok[RY], comp[RY], fluid[RY];
  unsigned cc[RY];
  float et[RY];
#pragma unroll
  for (int q = 0; q < RY; ++q) {
    const int j = j0 + ty + BYT * q;
    ok[q] = iok && j < g.ny;
    comp[q] = icomp && j <= g.nj2;
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
      const int o = (ty + BYT * q + 3) * LW + tx + 3;
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
    sTxy[hy.lds] = hyTxy;
    sTyy[hy.lds] = hyTyy;
    sTyz[hy.lds] = hyTyz;
    // RK operands of plane k
    float rr[RY], mo[RY][3], to[RY][3];
    bool fl[RY];
#pragma unroll
    for (int q = 0; q < RY; ++q) {
      const unsigned c = cc[q] + kz;
      rr[q] = comp[q] ? g.rrho[c] : 0.f;
      fl[q] = fluid[q] ? g.mu[c] <= SEIS_ZERO : false;
#pragma unroll
      for (int f = 0; f < 3; ++f) {
        const long e = f * fs + c;
        mo[q][f] = (comp[q] && KIND != 2) ? (KIND == 0 ? W[e] : g.m[e]) : 0.f;
        to[q][f] = (comp[q] && KIND != 0) ? g.t[e] : 0.f;
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
    __syncthreads();
    const float ze = g.zeta_z[k], rzk = g.rz[k];
#pragma unroll
    for (int q = 0; q < RY; ++q) {
      if (comp[q]) {
        const unsigned c = cc[q] + kz;
        const int o = (ty + BYT * q + 3) * LW + tx + 3;
        const float DxTxx = dlds<XF>(sTxx + o, 1) * xi;
        float DxTxy = dlds<XF>(sTxy + o, 1) * xi;
        float DxTxz = dlds<XF>(sTxz + o, 1) * xi;
        const float DyTyy = dlds<YF>(sTyy + o, LW) * et[q];
        float DyTxy = dlds<YF>(sTxy + o, LW) * et[q];
        float DyTyz = dlds<YF>(sTyz + o, LW) * et[q];
        const float DzTzz = dz5<ZF>(w[q][0]) * ze;
        float DzTxz = dz5<ZF>(w[q][1]) * ze;
        float DzTyz = dz5<ZF>(w[q][2]) * ze;
        if (fl[q]) {
          DxTxy = 0.f; DxTxz = 0.f; DyTxy = 0.f; DyTyz = 0.f; DzTxz = 0.f; DzTyz = 0.f;
        }
        const float Txx = cTxx[q], Tyy = cTyy[q], Txy = cTxy[q];
        const float Tzz = zc<ZF>(w[q][0]), Txz = zc<ZF>(w[q][1]), Tyz = zc<ZF>(w[q][2]);
        float h[3];
        h[0] = rr[q] * (DxTxx * rzk + DyTxy * rzk * rxs + DzTxz + (3.0f * Txz + Txx * xc - Tyy * xc) * rzk);
        h[1] = rr[q] * (DxTxy * rzk + DyTyy * rzk * rxs + DzTyz + (2.0f * Txy * xc + 3.0f * Tyz) * rzk);
        h[2] = rr[q] * (DxTxz * rzk + DyTyz * rzk * rxs + DzTzz + (2.0f * Tzz - Txx - Tyy + Txz * xc) * rzk);
#pragma unroll
        for (int f = 0; f < 3; ++f) {
          const long e = f * fs + c;
          if constexpr (KIND == 2) {
            g.wout[e] = to[q][f] + g.b * h[f];
          } else {
            g.wout[e] = mo[q][f] + g.a * h[f];
            g.t[e] = (KIND == 0 ? mo[q][f] : to[q][f]) + g.b * h[f];
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
template <int XF, int YF, int ZF, int KIND, int RY>
__global__ __launch_bounds__(NTH) void k_str(StageArgs g) {
  constexpr int TY = BYT * RY, LW = TX + 6, LH = TY + 6, PL = LW * LH;
  __shared__ float sh[2][3 * PL];  // Vx Vy Vz, double buffered
  const int tx = threadIdx.x, ty = threadIdx.y, tid = tx + TX * ty;
  int bx, by, bz;
  tile_of_block(g, bx, by, bz);
  const int i0 = g.ni1 + bx * TX, j0 = g.nj1 + by * TY;
  const int k0 = g.nk1 + bz * g.kchunk, k1 = min(k0 + g.kchunk - 1, g.nk2);
  const long fs = g.fs;
  const unsigned sz = (unsigned)g.sz;
  const float* __restrict__ W = g.win;
  const float* __restrict__ pV0 = W + VX * fs;
  const float* __restrict__ pV1 = W + VY * fs;
  const float* __restrict__ pV2 = W + VZ * fs;
  const int i = i0 + tx;
  const bool iok = i < g.nx, icomp = i <= g.ni2;
  const int ic = min(i, g.nx - 1);
  const float xi = g.xi_x[ic], rxs = g.rxsin[ic], xc = g.xcot[ic];
  const HaloSlot hx = xhalo_slot<XF, TY, LW>(g, i0, j0, tid);
  const HaloSlot hy = yhalo_slot<YF, TY, LW>(g, i0, j0, tid);

  float w[RY][3][5];
  bool ok[RY], comp[RY];
  unsigned cc[RY];
  float et[RY];
#pragma unroll
  for (int q = 0; q < RY; ++q) {
    const int j = j0 + ty + BYT * q;
    ok[q] = iok && j < g.ny;
    comp[q] = icomp && j <= g.nj2;
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
  unsigned kz = sz * (unsigned)k0;
  float hx0 = ld(pV0, hx, kz), hx1 = ld(pV1, hx, kz), hx2 = ld(pV2, hx, kz);
  float hy0 = ld(pV0, hy, kz), hy1 = ld(pV1, hy, kz), hy2 = ld(pV2, hy, kz);

  for (int k = k0; k <= k1; ++k, kz += sz) {
    float* s = sh[(k - k0) & 1];
#pragma unroll
    for (int q = 0; q < RY; ++q) {
      const int o = (ty + BYT * q + 3) * LW + tx + 3;
      s[o] = zc<ZF>(w[q][0]);
      s[PL + o] = zc<ZF>(w[q][1]);
      s[2 * PL + o] = zc<ZF>(w[q][2]);
    }
    if (hx.lds >= 0) {
      s[hx.lds] = hx0;
      s[PL + hx.lds] = hx1;
      s[2 * PL + hx.lds] = hx2;
    }
    s[hy.lds] = hy0;
    s[PL + hy.lds] = hy1;
    s[2 * PL + hy.lds] = hy2;
    const bool surf = g.freenode && k == g.nk2;
    // RK operands of plane k
    float la[RY], mu_[RY], mo[RY][6], to[RY][6];
#pragma unroll
    for (int q = 0; q < RY; ++q) {
      const unsigned c = cc[q] + kz;
      la[q] = comp[q] ? g.lam[c] : 0.f;
      mu_[q] = comp[q] ? g.mu[c] : 0.f;
#pragma unroll
      for (int f = 0; f < 6; ++f) {
        const long e = (TXX + f) * fs + c;
        float mv = (comp[q] && KIND != 2) ? (KIND == 0 ? W[e] : g.m[e]) : 0.f;
        if (surf && comp[q] && KIND != 2 && (f == 2 || f == 4 || f == 5))
          mv = g.msurf[(f == 2 ? 0 : (f == 4 ? 1 : 2)) * g.ps + cc[q]];
        mo[q][f] = mv;
        to[q][f] = (comp[q] && KIND != 0) ? g.t[e] : 0.f;
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
    __syncthreads();
    const float ze = g.zeta_z[k], rzk = g.rz[k];
#pragma unroll
    for (int q = 0; q < RY; ++q) {
      if (comp[q]) {
        const unsigned c = cc[q] + kz;
        const int o = (ty + BYT * q + 3) * LW + tx + 3;
        const float DxVx = dlds<XF>(s + o, 1) * xi;
        const float DxVy = dlds<XF>(s + PL + o, 1) * xi;
        const float DxVz = dlds<XF>(s + 2 * PL + o, 1) * xi;
        const float DyVx = dlds<YF>(s + o, LW) * et[q];
        const float DyVy = dlds<YF>(s + PL + o, LW) * et[q];
        const float DyVz = dlds<YF>(s + 2 * PL + o, LW) * et[q];
        float DzVx, DzVy, DzVz;
        if (g.freenode && k == g.nk2 - 1) {
          DzVx = dz22<ZF>(w[q][0]) * ze; DzVy = dz22<ZF>(w[q][1]) * ze; DzVz = dz22<ZF>(w[q][2]) * ze;
        } else if (g.freenode && k == g.nk2 - 2) {
          DzVx = dz24<ZF>(w[q][0]) * ze; DzVy = dz24<ZF>(w[q][1]) * ze; DzVz = dz24<ZF>(w[q][2]) * ze;
        } else {
          DzVx = dz5<ZF>(w[q][0]) * ze; DzVy = dz5<ZF>(w[q][1]) * ze; DzVz = dz5<ZF>(w[q][2]) * ze;
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
#pragma unroll
        for (int f = 0; f < 6; ++f) {
          const long e = (TXX + f) * fs + c;
          if constexpr (KIND == 2) {
            g.wout[e] = to[q][f] + g.b * h[f];
          } else {
            g.wout[e] = mo[q][f] + g.a * h[f];
            g.t[e] = (KIND == 0 ? mo[q][f] : to[q][f]) + g.b * h[f];
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

}  // namespace fw
