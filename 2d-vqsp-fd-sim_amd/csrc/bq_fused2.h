// bq_fused2.h - two time steps per sweep of the interior (temporal blocking).
//
// The single-step fused tile (fused_block) moves 8 fields in, 8 out and one
// material index per cell and step. A two-step tile takes its cells from
// level n (set A) to level n+2 (set C) in one march down its rows, so the same
// bytes serve two steps. Level n+1 lives only in LDS and registers, except for
// the cells within two of the region's edge (the rim), which are also written
// to the scratch set T: the boundary cells outside the region are advanced one
// step at a time by the per-cell and run kernels (A -> T, then T -> C), and
// their stencils reach two cells into the region.
//
// Per march row s a block forms, over its W own columns plus halo columns,
//   level n+1 stresses  (S1) of row s     at columns c-6 .. c+W+5
//   level n+1 velocities(V1) of row s-2   at columns c-4 .. c+W+3
//   level n+2 stresses  (S2) of row s-4   at columns c-2 .. c+W+1
//   level n+2 velocities(V2) of row s-6   at its own columns c .. c+W-1
// from level n velocities loaded at columns c-8 .. c+W+7. Every value is the
// interior formula of ring_stress / ring_velocity, evaluated exactly as the
// single-step kernels do, so the result is bit-identical to two single steps;
// the region is chosen so that every cell of these halos uses the interior
// formulas (columns [4, MX-2], stress rows >= 3, velocity rows >= 2, rows
// <= MZ'-2), and the source term enters S1 with step itile and S2 with itile+1.
//
// Thread t owns column c+t. Halo duties, all in the first wave:
//   t < 16       loads the level-n velocities of one extra column
//   t < 12       forms S1 at one extra column (c-6 .. c-1, c+W .. c+W+5)
//   2 <= t < 10  also V1 there (c-4 .. c-1, c+W .. c+W+3)
//   4 <= t < 8   also S2 there (c-2, c-1, c+W, c+W+1)
// LDS column j holds global column c-8+j.
#pragma once
#include <hip/hip_runtime.h>

#include "bq_kernels.h"

namespace bqk {

using bq::Fused2;

// The scratch set T (level n+1 of the rim cells): the 8 fields.
struct F8 {
  float *f[8];  // U W QX QZ TXX TZZ TXZ PRES
};

constexpr int kF2W = 256;            // own columns of a two-step tile (threads per block)
constexpr int kF2T = kF2W;           // threads of a two-step block
constexpr int kF2R = 256;            // rows of a two-step tile
constexpr int kF2L = kF2W + 16;      // LDS row: columns c-8 .. c+kF2W+7

struct Fused2Lds {
  float sU[5][kF2L], sW[6][kF2L], sZ[6][kF2L], sQ[4][kF2L];   // level n velocities
  float aX[3][kF2L], aXZ[3][kF2L], aP[3][kF2L];               // level n+1 TXX, TXZ, PRES
  float vU[5][kF2L], vW[6][kF2L], vZ[6][kF2L], vQ[6][kF2L];   // level n+1 velocities
  float bX[3][kF2L], bXZ[3][kF2L], bP[3][kF2L];               // level n+2 TXX, TXZ, PRES
};

// Level n+1 stress history of one column: rows s-4 .. s (index 4 = row s).
struct S1Hist {
  float xx[5], zz[5], xz[5], pr[5];
  __device__ __forceinline__ void push(const float *o) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      xx[r] = xx[r + 1];
      zz[r] = zz[r + 1];
      xz[r] = xz[r + 1];
      pr[r] = pr[r + 1];
    }
    xx[4] = o[0];
    zz[4] = o[1];
    xz[4] = o[2];
    pr[4] = o[3];
  }
};

__device__ __forceinline__ bool f2_rim(const Fused2 &fu, int i, int l) {
  return i < fu.x0 + 2 || i >= fu.x1 - 2 || l < fu.z0 + 2 || l >= fu.z1 - 2;
}

// Block b of the two-step region. k: current set A in U0.., set C in U1..;
// rim: set T. itile: the first of the two steps.
__device__ __forceinline__ void fused2_block(const KParams &k, const Fused2 &fu, const F8 &rim,
                                             int b, int itile, Fused2Lds &L) {
  constexpr int W = kF2W, LW = kF2L;
  const int t = threadIdx.x;
  const int bx = b % fu.nfx, bz = b / fu.nfx;
  const int c = fu.x0 + bx * W;
  const int a = fu.z0 + bz * fu.fr;
  const int e = min(a + fu.fr, fu.z1);
  const int P = k.P, OFF = k.OFF;
  const int i = c + t, gx = i + OFF, q = t + 8;
  // extra columns
  const bool hl = t < 16;
  const int hlq = t < 8 ? t : W + t;              // loaded column c-8+hlq
  const int hlx = c - 8 + hlq + OFF;
  const bool h1 = t < 12;
  const int h1q = t < 6 ? t + 2 : W + 2 + t;      // S1 column c-8+h1q
  const int h1i = c - 8 + h1q;
  const int h1x = h1i + OFF;
  const bool h2 = t >= 2 && t < 10;
  const bool h3 = t >= 4 && t < 8;
  const float *__restrict__ U = k.U0;
  const float *__restrict__ Wf = k.W0;
  const float *__restrict__ QX = k.QX0;
  const float *__restrict__ QZ = k.QZ0;
  const int s0 = a - 6, s1 = e + 5;  // march rows
  // prefill: U rows s0-2 .. s0, W and QZ rows s0-1 .. s0+1
  for (int r = s0 - 2; r <= s0; ++r) {
    L.sU[r % 5][q] = U[r * P + gx];
    L.sW[(r + 1) % 6][q] = Wf[(r + 1) * P + gx];
    L.sZ[(r + 1) % 6][q] = QZ[(r + 1) * P + gx];
    if (hl) {
      L.sU[r % 5][hlq] = U[r * P + hlx];
      L.sW[(r + 1) % 6][hlq] = Wf[(r + 1) * P + hlx];
      L.sZ[(r + 1) % 6][hlq] = QZ[(r + 1) * P + hlx];
    }
  }
  S1Hist ho{}, hh{};                 // level n+1 stress history: own column, extra column
  float bxz[4] = {0, 0, 0, 0};       // level n+2 TXZ, rows s-7 .. s-4 (own column)
  float bzz[5] = {0, 0, 0, 0, 0};    // level n+2 TZZ, rows s-8 .. s-4
  float bpp[5] = {0, 0, 0, 0, 0};    // level n+2 PRES, rows s-8 .. s-4
  int jo[7] = {0, 0, 0, 0, 0, 0, 0};  // material of rows s .. s-6 (own column)
  int jh[5] = {0, 0, 0, 0, 0};        // material of rows s .. s-4 (extra column)
  // row s+1's values, loaded while row s is computed
  float pu, pw, pz, pq, fhu = 0, fhw = 0, fhz = 0, fhq = 0;
  float ftxx, ftzz, ftxz, fpr, fhtxx = 0, fhtzz = 0, fhtxz = 0, fhpr = 0;
  int fjm, fhjm = 0;
  auto fetch = [&](int s) {
    const int r = s * P;
    // own column: streaming hint, as in the single-step tiles
    pu = stream_load(&U[r + P + gx]); pw = stream_load(&Wf[r + 2 * P + gx]);
    pz = stream_load(&QZ[r + 2 * P + gx]); pq = stream_load(&QX[r + gx]);
    if (hl) { fhu = U[r + P + hlx]; fhw = Wf[r + 2 * P + hlx]; fhz = QZ[r + 2 * P + hlx]; fhq = QX[r + hlx]; }
    ftxx = stream_load(&k.TXX0[r + gx]); ftzz = stream_load(&k.TZZ0[r + gx]);
    ftxz = stream_load(&k.TXZ0[r + gx]); fpr = stream_load(&k.PR0[r + gx]);
    fjm = jm_at(k, r + gx);
    if (h1) {
      fhtxx = k.TXX0[r + h1x]; fhtzz = k.TZZ0[r + h1x]; fhtxz = k.TXZ0[r + h1x];
      fhpr = k.PR0[r + h1x]; fhjm = jm_at(k, r + h1x);
    }
  };
  fetch(s0);
  for (int s = s0; s <= s1; ++s) {
    // ---- level n velocities entering the rings; old stresses and index of row s
    L.sU[(s + 1) % 5][q] = pu;
    L.sW[(s + 2) % 6][q] = pw;
    L.sZ[(s + 2) % 6][q] = pz;
    L.sQ[s & 3][q] = pq;
    if (hl) {
      L.sU[(s + 1) % 5][hlq] = fhu;
      L.sW[(s + 2) % 6][hlq] = fhw;
      L.sZ[(s + 2) % 6][hlq] = fhz;
      L.sQ[s & 3][hlq] = fhq;
    }
    const float otxx = ftxx, otzz = ftzz, otxz = ftxz, opr = fpr;
    const float htxx = fhtxx, htzz = fhtzz, htxz = fhtxz, hpr = fhpr;
#pragma unroll
    for (int r = 6; r > 0; --r) jo[r] = jo[r - 1];
    jo[0] = fjm;
#pragma unroll
    for (int r = 4; r > 0; --r) jh[r] = jh[r - 1];
    jh[0] = fhjm;
    if (s < s1) fetch(s + 1);
    __syncthreads();
    // ---- level n+1 stresses of row s
    const int ns = s % 3;
    {
      float o[4];
      ring_stress<5, 6, 6, 4, LW>(k, L.sU, L.sW, L.sZ, L.sQ, s, q, i, otxx, otzz, otxz, opr, jo[0],
                                  itile, o);
      L.aX[ns][q] = o[0];
      L.aXZ[ns][q] = o[2];
      L.aP[ns][q] = o[3];
      ho.push(o);
      if (s >= a && s < e && f2_rim(fu, i, s)) {
        const int r = s * P + gx;
        rim.f[4][r] = o[0];
        rim.f[5][r] = o[1];
        rim.f[6][r] = o[2];
        rim.f[7][r] = o[3];
      }
    }
    if (h1) {
      float o[4];
      ring_stress<5, 6, 6, 4, LW>(k, L.sU, L.sW, L.sZ, L.sQ, s, h1q, h1i, htxx, htzz, htxz, hpr,
                                  jh[0], itile, o);
      L.aX[ns][h1q] = o[0];
      L.aXZ[ns][h1q] = o[2];
      L.aP[ns][h1q] = o[3];
      hh.push(o);
    }
    __syncthreads();
    // ---- level n+1 velocities of row l1 = s-2 (rows a-4 .. e+3)
    const int l1 = s - 2;
    if (l1 >= a - 4) {
      const int sl = l1 % 3;
      {
        float v[4];
        ring_velocity(k, &L.aX[sl][q - 2], &L.aXZ[sl][q - 2], &L.aP[sl][q - 2], ho.xz[1], ho.xz[2],
                      ho.xz[3], ho.xz[4], ho.zz[0], ho.zz[1], ho.zz[2], ho.zz[3], ho.pr[0],
                      ho.pr[1], ho.pr[2], ho.pr[3], L.sQ[l1 & 3][q], L.sU[l1 % 5][q],
                      L.sZ[l1 % 6][q], L.sW[l1 % 6][q], jo[2], v);
        L.vQ[l1 % 6][q] = v[0];
        L.vU[l1 % 5][q] = v[1];
        L.vZ[l1 % 6][q] = v[2];
        L.vW[l1 % 6][q] = v[3];
        if (l1 >= a && l1 < e && f2_rim(fu, i, l1)) {
          const int r = l1 * P + gx;
          rim.f[2][r] = v[0];
          rim.f[0][r] = v[1];
          rim.f[3][r] = v[2];
          rim.f[1][r] = v[3];
        }
      }
      if (h2) {
        float v[4];
        ring_velocity(k, &L.aX[sl][h1q - 2], &L.aXZ[sl][h1q - 2], &L.aP[sl][h1q - 2], hh.xz[1],
                      hh.xz[2], hh.xz[3], hh.xz[4], hh.zz[0], hh.zz[1], hh.zz[2], hh.zz[3],
                      hh.pr[0], hh.pr[1], hh.pr[2], hh.pr[3], L.sQ[l1 & 3][h1q], L.sU[l1 % 5][h1q],
                      L.sZ[l1 % 6][h1q], L.sW[l1 % 6][h1q], jh[2], v);
        L.vQ[l1 % 6][h1q] = v[0];
        L.vU[l1 % 5][h1q] = v[1];
        L.vZ[l1 % 6][h1q] = v[2];
        L.vW[l1 % 6][h1q] = v[3];
      }
    }
    __syncthreads();
    // ---- level n+2 stresses of row r2 = s-4 (rows a-2 .. e+1); the old
    //      stresses are level n+1 of that row (history index 0)
    const int r2 = s - 4;
    if (r2 >= a - 2) {
      const int nb = r2 % 3;
      {
        float o[4];
        ring_stress<5, 6, 6, 6, LW>(k, L.vU, L.vW, L.vZ, L.vQ, r2, q, i, ho.xx[0], ho.zz[0],
                                    ho.xz[0], ho.pr[0], jo[4], itile + 1, o);
        L.bX[nb][q] = o[0];
        L.bXZ[nb][q] = o[2];
        L.bP[nb][q] = o[3];
        bxz[0] = bxz[1]; bxz[1] = bxz[2]; bxz[2] = bxz[3]; bxz[3] = o[2];
        bzz[0] = bzz[1]; bzz[1] = bzz[2]; bzz[2] = bzz[3]; bzz[3] = bzz[4]; bzz[4] = o[1];
        bpp[0] = bpp[1]; bpp[1] = bpp[2]; bpp[2] = bpp[3]; bpp[3] = bpp[4]; bpp[4] = o[3];
        if (r2 >= a && r2 < e) {
          const int r = r2 * P + gx;
          stream_store(&k.TXX1[r], o[0]);
          stream_store(&k.TZZ1[r], o[1]);
          stream_store(&k.TXZ1[r], o[2]);
          stream_store(&k.PR1[r], o[3]);
        }
      }
      if (h3) {
        float o[4];
        ring_stress<5, 6, 6, 6, LW>(k, L.vU, L.vW, L.vZ, L.vQ, r2, h1q, h1i, hh.xx[0], hh.zz[0],
                                    hh.xz[0], hh.pr[0], jh[4], itile + 1, o);
        L.bX[nb][h1q] = o[0];
        L.bXZ[nb][h1q] = o[2];
        L.bP[nb][h1q] = o[3];
      }
    }
    __syncthreads();
    // ---- level n+2 velocities of row l2 = s-6 (own rows a .. e-1)
    const int l2 = s - 6;
    if (l2 >= a) {
      const int sl = l2 % 3;
      float v[4];
      ring_velocity(k, &L.bX[sl][q - 2], &L.bXZ[sl][q - 2], &L.bP[sl][q - 2], bxz[0], bxz[1],
                    bxz[2], bxz[3], bzz[0], bzz[1], bzz[2], bzz[3], bpp[0], bpp[1], bpp[2], bpp[3],
                    L.vQ[l2 % 6][q], L.vU[l2 % 5][q], L.vZ[l2 % 6][q], L.vW[l2 % 6][q], jo[6], v);
      const int r = l2 * P + gx;
      stream_store(&k.QX1[r], v[0]);
      stream_store(&k.U1[r], v[1]);
      stream_store(&k.QZ1[r], v[2]);
      stream_store(&k.W1[r], v[3]);
    }
  }
}

}  // namespace bqk
