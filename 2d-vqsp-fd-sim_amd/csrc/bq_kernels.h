// bq_kernels.h - device code of one time step of the Biot time loop for
// gfx950, shared by the single-grid launches (bq_device.hip) and the batched
// launches over many grids (bq_batch.hip).
//
// Each cell is evaluated with the reference's operation order and no
// floating-point contraction (the sources are built with -ffp-contract=off),
// so the fields are bit-identical to the source-order binary32 evaluation
// whichever launch (per-cell blocks, wavefront runs, fused tiles) computes
// them.
#pragma once
#include <hip/hip_runtime.h>

#include "bq_internal.h"

using bq::Fused;
using bq::GridDesc;
using bq::Split;
using bq::kMaxCells;
using bq::kMaxRuns;
using bq::kEdgeCols;

// Named (not anonymous) so that the host structures holding KParams have the
// same type in every source file that includes this header.
namespace bqk {


constexpr int kEdgeBlock = 256;

struct KParams {
  int MX, MZ, T, PMLL, PMLZH, NX, P, OFF, LIN_X, LIN_Z, MR;
  int SI;                // stress imaging (else explicit surface rows 0 and 1)
  int AFDA_ZZ, AFDA_XZ;  // one-sided surface options of the explicit surface
  int SECOND;            // second-order build: A = 0, B = 1
  float DT, A, B, AH, BH, CH, BC;
  // fields of the current step (read) and of the next step (written); the two
  // sets trade places after every step
  const float *U0, *W0, *QX0, *QZ0, *TXX0, *TZZ0, *TXZ0, *PR0;
  float *U1, *W1, *QX1, *QZ1, *TXX1, *TZZ1, *TXZ1, *PR1;
  // material index (1-based), 16-bit where the table has at most 65535
  // entries (JM16), else 32-bit (JM); the other pointer is null
  const int *JM;
  const unsigned short *JM16;
  const float4 *stab;  // per material: {MU,SXX1,SXX2,SXX3},{SZZ1,SZZ2,SZZ3,P1},{P2,P3,-,-}
  const float4 *vtab;  // per material: {eQX,R1QX,R2QX,fU},{R1U,R3U,eQZ,R1QZ},{R2QZ,fW,R1W,R3W}
  const float2 *atab;  // per material: {AUXIL_W, AUXIL_QZ}
  const float *x1[8], *x2[8], *z1[4], *z2[4];
  float *PZXL, *PZXH, *PXXL, *PXXH, *OXXL, *OXXH, *PXZH, *PZZH, *OZZH;
  float *RXXL, *RXXH, *SXXL, *SXXH, *GXXL, *GXXH, *HXXL, *HXXH;
  float *RZXL, *RZXH, *GZXL, *GZXH;
  float *RXZH, *GXZH, *RZZH, *SZZH, *GZZH, *HZZH;
  float *SZc, *SZp;     // bottom history: [4 fields][NX][3] current / previous
  const float2 *dsrc;   // per step: ((S-Sprev)/DT, (F-Fprev)/DT)
  const int4 *rec;      // receivers: irecx, lrecx, irecz, lrecz
  float *traces;        // [MT2][MR][6]
  float abw[8], abh[8];
};

#define FA(a, i, l) (a)[(size_t)(l) * (size_t)k.P + (size_t)((i) + k.OFF)]
#define XL1(t, j, l) (k.x1[t][(l) * (T + 1) + (j)])
#define XL2(t, j, l) (k.x2[t][(l) * (T + 2) + (j)])
// the bottom PML coefficients sit as [j][ix] on the device (the reference's
// (0:T, NXc) arrays transposed), so that a wavefront's cells of one row read
// consecutive addresses
#define ZH1(t, j, ix) (k.z1[t][(size_t)(j) * k.NX + (ix)])
#define ZH2(t, j, ix) (k.z2[t][(size_t)(j) * k.NX + (ix)])

// Loads and stores with the streaming (non-temporal) hint, for the field rows
// the fused tiles read in their own columns and the new rows they write (16
// values per cell, about 2 GB per step on an 8040 x 4023 grid), which no block
// reads again in the launch. L2 hits and misses stay as they were; what
// changes is below L2: the streamed rows no longer fill the last-level cache,
// which then keeps the material tables and the halo rows and columns that L2
// misses reach. Measured on that grid, k_step:
// 519 us with plain loads and stores, 493 us with streaming stores, 463-466 us
// with streaming stores and own-column loads; the hint on the halo-column and
// material-index loads as well made it 478 us, on the own material index
// alone 471 us.
__device__ __forceinline__ void stream_store(float *p, float v) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ float stream_load(const float *p) { return __builtin_nontemporal_load(p); }

// Material index of cell c (row-major with the field pitch), 0-based.
__device__ __forceinline__ int jm_at(const KParams &k, size_t c) {
  return (k.JM16 ? (int)k.JM16[c] : k.JM[c]) - 1;
}

// One-sided surface stencils of the explicit free surface (binary32 quotients,
// as the reference's constant expressions fold them).
constexpr float c17_24 = 17.0f / 24.0f, c3_8 = 3.0f / 8.0f, c5_24 = 5.0f / 24.0f;
constexpr float c1_24 = 1.0f / 24.0f, c11_12 = 11.0f / 12.0f;
constexpr float c35_8 = 35.0f / 8.0f, c35_24 = 35.0f / 24.0f, c21_40 = 21.0f / 40.0f;
constexpr float c5_56 = 5.0f / 56.0f, c31_24 = 31.0f / 24.0f, c29_24 = 29.0f / 24.0f;
constexpr float c3_40 = 3.0f / 40.0f, c1_168 = 1.0f / 168.0f, c1_3 = 1.0f / 3.0f;
constexpr float c1_22 = 1.0f / 22.0f, c577_528 = 577.0f / 528.0f, c201_176 = 201.0f / 176.0f;
constexpr float c9_176 = 9.0f / 176.0f, c1_528 = 1.0f / 528.0f;

// ------------------------------------------------------------------ stress
// One cell (i,l) of the stress half-step, l in [2, PMLZH]:
//   l in [2, MZ'-2]  HPLANE_IC_STRESS / HPLANE_LIN_STRESS (hplane_ic_stress.F90:48-117)
//   l = MZ'-1        HPLANE_MZ1_STRESS (hplane_mz1_stress.F90:50-130)
//   l in [MZ', PMLZH-1]  HPLANE_PML_STRESS, LL = l-MZ'+1 (hplane_pml_stress.F90:53-134)
//   l = PMLZH        HPLANE_PMLZH0_STRESS, TXZ only (hplane_pmlzh0_stress.F90:43-68)
// and the stress-imaging rows (hplane_0_stress.F90:47-51, hplane_1_stress.F90:44-49).
// Without stress imaging also l = 0 and 1 (HPLANE_0_STRESS, HPLANE_1_STRESS,
// hplane_0_stress.F90:57-119, hplane_1_stress.F90:51-126): EXZ = 0 on row 0,
// STRAIN_UW_XZ_1 on row 1 (strain_uw_xz_1.F90:62-114), STRAIN_UW_ZZ_0 /
// STRAIN_Q_ZZ_0 on row 0 (strain_uw_zz_0.F90:45-62, strain_q_zz_0.F90:40-57).
__device__ __forceinline__ void stress_cell(const KParams &k, const int i, const int l,
                                            const int itile) {
  const int T = k.T, MX = k.MX, MZ = k.MZ, PMLZH = k.PMLZH, NX = k.NX;
  const int ix = i - k.PMLL;
  const bool zh0 = (l == PMLZH);
  const bool pmlrow = (l >= MZ && l < PMLZH);
  const bool zpml = zh0 || pmlrow;
  const int LL = zh0 ? T + 1 : (pmlrow ? l - MZ + 1 : 0);
  const float C = k.CH, DT = k.DT;
  const int jm = jm_at(k, (size_t)l * k.P + ix);
  const float4 s0 = k.stab[3 * jm];

  // ---- shear strain and TXZ: STRAIN_UW_XZ (strain_uw_xz.F90:61-103) or
  //      STRAIN_UW_XZ_2ND_PML (strain_uw_xz_2nd_pml.F90:69-121), STRESS_XZ
  float txz = FA(k.TXZ0, i, l);
  if (i >= 2 - T && i <= MX + T) {
    float exz;
    if (l == 0) {
      exz = 0.0f;
    } else if (!zpml) {
      if (i <= 2) {
        const float dzxa = C * (FA(k.W0, i, l) - FA(k.W0, i - 1, l));
        float &pm = k.PZXL[(size_t)l * (T + 2) + ix];
        const int j = 3 - i;
        exz = C * (FA(k.U0, i, l) - FA(k.U0, i, l - 1)) + XL2(0, j, l) * (XL2(3, j, l) * dzxa + pm);
        pm = XL2(1, j, l) * pm + XL2(2, j, l) * dzxa;
      } else if (i <= MX - 1) {
        if (l != 1) {
          exz = k.AH * (FA(k.U0, i, l + 1) - FA(k.U0, i, l - 2) + FA(k.W0, i + 1, l) - FA(k.W0, i - 2, l)) +
                k.BH * (FA(k.U0, i, l) - FA(k.U0, i, l - 1) + FA(k.W0, i, l) - FA(k.W0, i - 1, l));
        } else if (!k.AFDA_XZ) {
          exz = k.AH * (FA(k.W0, i + 1, 1) - FA(k.W0, i - 2, 1)) + k.BH * (FA(k.W0, i, 1) - FA(k.W0, i - 1, 1)) +
                C * (FA(k.U0, i, 1) - FA(k.U0, i, 0));
        } else {
          const float duikm1 = -(k.AH * (FA(k.W0, i + 1, 0) - FA(k.W0, i - 2, 0))) -
                               k.BH * (FA(k.W0, i, 0) - FA(k.W0, i - 1, 0));
          exz = k.AH * (FA(k.W0, i + 1, 1) - FA(k.W0, i - 2, 1)) + k.BH * (FA(k.W0, i, 1) - FA(k.W0, i - 1, 1)) -
                c1_22 * duikm1 +
                C * (-(c577_528 * FA(k.U0, i, 0)) + c201_176 * FA(k.U0, i, 1) - c9_176 * FA(k.U0, i, 2) +
                     c1_528 * FA(k.U0, i, 3));
        }
      } else {
        const float dzxa = C * (FA(k.W0, i, l) - FA(k.W0, i - 1, l));
        float &pm = k.PZXH[(size_t)l * (T + 1) + (i - MX)];
        const int j = i - (MX - 1);
        exz = C * (FA(k.U0, i, l) - FA(k.U0, i, l - 1)) + XL2(4, j, l) * (XL2(7, j, l) * dzxa + pm);
        pm = XL2(5, j, l) * pm + XL2(6, j, l) * dzxa;
      }
    } else {
      const float dzxa = C * (FA(k.W0, i, l) - FA(k.W0, i - 1, l));
      const float dxza = C * (FA(k.U0, i, l) - FA(k.U0, i, l - 1));
      float &pz = k.PXZH[(size_t)LL * NX + ix];
      const float zp = ZH2(0, LL, ix) * (ZH2(3, LL, ix) * dxza + pz);
      if (i <= 2) {
        float &pm = k.PZXL[(size_t)l * (T + 2) + ix];
        const int j = 3 - i;
        exz = zp + XL2(0, j, l) * (XL2(3, j, l) * dzxa + pm);
        pm = XL2(1, j, l) * pm + XL2(2, j, l) * dzxa;
      } else if (i <= MX - 1) {
        exz = zp + dzxa;
      } else {
        float &pm = k.PZXH[(size_t)l * (T + 1) + (i - MX)];
        const int j = i - (MX - 1);
        exz = zp + XL2(4, j, l) * (XL2(7, j, l) * dzxa + pm);
        pm = XL2(5, j, l) * pm + XL2(6, j, l) * dzxa;
      }
      pz = ZH2(1, LL, ix) * pz + ZH2(2, LL, ix) * dxza;
    }
    txz = txz + DT * (s0.x * exz);
  }
  // TXZ(:,2) = 0 after the images are taken (hplane_1_stress.F90:48)
  FA(k.TXZ1, i, l) = (k.SI && l == 2) ? 0.0f : txz;
  if (k.SI && l == 3) FA(k.TXZ1, i, 1) = -txz;
  if (k.SI && l == 4) FA(k.TXZ1, i, 0) = -txz;
  if (zh0) return;

  // ---- normal strains and TXX, TZZ, PRES: STRAIN_UW_XX / STRAIN_Q_XX
  //      (strain_uw_xx.F90:60-107), STRAIN_UW_ZZ / STRAIN_Q_ZZ (strain_uw_zz.F90:49-60)
  //      or their _2ND_PML forms (strain_uw_zz_2nd_pml.F90:55-65), STRESS_NN /
  //      STRESS_LIN_NN (stress_nn.F90:49-61, stress_lin_nn.F90:49-70)
  float tzz = FA(k.TZZ0, i, l), pr = FA(k.PR0, i, l);
  if (i >= 2 - T && i <= MX - 1 + T) {
    float exx, qxx, ezz, qzz;
    if (i <= 1) {
      const int j = 2 - i;
      const size_t mo = (size_t)l * (T + 1) + ix;
      const float a = XL1(0, j, l), b = XL1(1, j, l), c = XL1(2, j, l), d = XL1(3, j, l);
      const float du = C * (FA(k.U0, i + 1, l) - FA(k.U0, i, l));
      const float dq = C * (FA(k.QX0, i + 1, l) - FA(k.QX0, i, l));
      float pu = k.PXXL[mo], pq = k.OXXL[mo];
      exx = a * (d * du + pu);
      qxx = a * (d * dq + pq);
      k.PXXL[mo] = b * pu + c * du;
      k.OXXL[mo] = b * pq + c * dq;
    } else if (i == 2 || i == MX - 1) {
      exx = C * (FA(k.U0, i + 1, l) - FA(k.U0, i, l));
      qxx = C * (FA(k.QX0, i + 1, l) - FA(k.QX0, i, l));
    } else if (i <= MX - 2) {
      exx = k.AH * (FA(k.U0, i + 2, l) - FA(k.U0, i - 1, l)) + k.BH * (FA(k.U0, i + 1, l) - FA(k.U0, i, l));
      qxx = k.AH * (FA(k.QX0, i + 2, l) - FA(k.QX0, i - 1, l)) + k.BH * (FA(k.QX0, i + 1, l) - FA(k.QX0, i, l));
    } else {
      const int j = i - (MX - 1);
      const size_t mo = (size_t)l * (T + 1) + (i - MX);
      const float a = XL1(4, j, l), b = XL1(5, j, l), c = XL1(6, j, l), d = XL1(7, j, l);
      const float du = C * (FA(k.U0, i + 1, l) - FA(k.U0, i, l));
      const float dq = C * (FA(k.QX0, i + 1, l) - FA(k.QX0, i, l));
      float pu = k.PXXH[mo], pq = k.OXXH[mo];
      exx = a * (d * du + pu);
      qxx = a * (d * dq + pq);
      k.PXXH[mo] = b * pu + c * du;
      k.OXXH[mo] = b * pq + c * dq;
    }
    if (l == 0) {
      ezz = (C * (FA(k.W0, i, 1) - FA(k.W0, i, 0)));
      qzz = (C * (FA(k.QZ0, i, 1) - FA(k.QZ0, i, 0)));
      if (k.AFDA_ZZ && i >= 2 && i <= MX - 1) {
        ezz = C * (-(c11_12 * FA(k.W0, i, 0)) + c17_24 * FA(k.W0, i, 1) + c3_8 * FA(k.W0, i, 2) -
                   c5_24 * FA(k.W0, i, 3) + c1_24 * FA(k.W0, i, 4));
        qzz = C * (-(c11_12 * FA(k.QZ0, i, 0)) + c17_24 * FA(k.QZ0, i, 1) + c3_8 * FA(k.QZ0, i, 2) -
                   c5_24 * FA(k.QZ0, i, 3) + c1_24 * FA(k.QZ0, i, 4));
      }
    } else if (!pmlrow && l != MZ - 1) {
      ezz = (C * (FA(k.W0, i, l + 1) - FA(k.W0, i, l)));
      qzz = (C * (FA(k.QZ0, i, l + 1) - FA(k.QZ0, i, l)));
      if (i >= 2 && i <= MX - 1) {
        ezz = k.AH * (FA(k.W0, i, l + 2) - FA(k.W0, i, l - 1)) + k.BC * ezz;
        qzz = k.AH * (FA(k.QZ0, i, l + 2) - FA(k.QZ0, i, l - 1)) + k.BC * qzz;
      }
    } else {
      const float a = ZH1(0, LL, ix), b = ZH1(1, LL, ix), c = ZH1(2, LL, ix), d = ZH1(3, LL, ix);
      const float dw = C * (FA(k.W0, i, l + 1) - FA(k.W0, i, l));
      const float dq = C * (FA(k.QZ0, i, l + 1) - FA(k.QZ0, i, l));
      const size_t mo = (size_t)LL * NX + ix;
      float pw = k.PZZH[mo], pq = k.OZZH[mo];
      ezz = a * (d * dw + pw);
      qzz = a * (d * dq + pq);
      k.PZZH[mo] = b * pw + c * dw;
      k.OZZH[mo] = b * pq + c * dq;
    }
    const float4 s1 = k.stab[3 * jm + 1];
    const float4 s2 = k.stab[3 * jm + 2];
    const float rxx = s0.y * exx + s0.z * ezz + s0.w * qxx + s0.w * qzz;
    const float rzz = s1.x * exx + s1.y * ezz + s1.z * qxx + s1.z * qzz;
    const float rpp = s1.w * exx + s2.x * ezz + s2.y * qxx + s2.y * qzz;
    float txx = FA(k.TXX0, i, l);
    if (l == k.LIN_Z && i == k.LIN_X) {
      const float2 ds = k.dsrc[itile - 1];
      txx = txx + DT * (rxx - ds.x);
      tzz = tzz + DT * (rzz - ds.x);
      pr = pr + DT * (rpp + ds.y);
    } else {
      txx = txx + DT * rxx;
      tzz = tzz + DT * rzz;
      pr = pr + DT * rpp;
    }
    FA(k.TXX1, i, l) = txx;
    FA(k.TZZ1, i, l) = tzz;
    FA(k.PR1, i, l) = pr;
  }
  if (k.SI && l == 2) { FA(k.TZZ1, i, 1) = -tzz; FA(k.PR1, i, 1) = -pr; }
  if (k.SI && l == 3) { FA(k.TZZ1, i, 0) = -tzz; FA(k.PR1, i, 0) = -pr; }
}

// ---------------------------------------------------------------- velocity
// One cell (i,l) of the velocity half-step, l in [2, PMLZH-1]:
//   l in [2, MZ'-2]        HPLANE_IC_VELOCITY (hplane_ic_velocity.F90:59-202):
//                          INH_QX_INT, INH_V_UV_INT (inh_qx_int.F90:40-54,
//                          inh_v_uv_int.F90:41-56) inside, INH_QX_2ND /
//                          INH_V_U_2ND (inh_qx_2nd.F90:54-103, inh_v_u_2nd.F90:56-104)
//                          in the side PML; likewise QZ/W
//   l in [MZ'-1, PMLZH-1]  HPLANE_PML_VELOCITY, LL = l-MZ'+1 (hplane_pml_velocity.F90:64-180)
// and without stress imaging l = 0 (HPLANE_0_VELOCITY, hplane_0_velocity.F90:62-179:
// INH_QX_INT_0, INH_V_UV_INT_0, INH_QZ_INT_0, INH_V_W_INT_0, INH_*_2ND_0) and
// l = 1 (HPLANE_1_VELOCITY, hplane_1_velocity.F90:79-207: the row-2 forms except
// INH_V_QZ_INT_1 / INH_V_W_INT_1, inh_v_qz_int_1.F90:49-79, inh_v_w_int_1.F90:51-82).
// The new QX, U, QZ, W of cell (i,l) into o[0..3], PML memories updated in
// place when ST (a recompute of a neighbour's cell leaves them alone). OUT:
// o also holds, for a field the cell does not update, its value in the set
// being written (which no kernel writes this step).
template <bool ST, bool OUT>
__device__ __forceinline__ void velocity_cell_t(const KParams &k, const int i, const int l, float *o) {
  const int T = k.T, MX = k.MX, NX = k.NX;
  const int ix = i - k.PMLL;
  const bool pmlrow = (l >= k.MZ - 1);
  const int LL = l - k.MZ + 1;
  const float A = k.A, B = k.B;
  const int jm = jm_at(k, (size_t)l * k.P + ix);
  const float4 v0 = k.vtab[3 * jm], v1 = k.vtab[3 * jm + 1], v2 = k.vtab[3 * jm + 2];
  // v0 = {eQX, R1QX, R2QX, fU}, v1 = {R1U, R3U, eQZ, R1QZ}, v2 = {R2QZ, fW, R1W, R3W}

  // ---- QX and U, x in [3-T, MX-1+T]
  if (i >= 3 - T && i <= MX - 1 + T) {
    const float q = FA(k.QX0, i, l);
    const float u = FA(k.U0, i, l);
    float qn, un;
    const bool inner = (i >= 4 && i <= MX - 2);
    if (!pmlrow) {
      if (inner) {
        float sx;
        if (l != 0) {
          sx = A * (FA(k.TXX1, i + 1, l) - FA(k.TXX1, i - 2, l) + FA(k.TXZ1, i, l + 2) - FA(k.TXZ1, i, l - 1)) +
               B * (FA(k.TXX1, i, l) - FA(k.TXX1, i - 1, l) + FA(k.TXZ1, i, l + 1) - FA(k.TXZ1, i, l));
        } else if (!k.SECOND) {  // TXZ(i,0) = 0 on the surface: one-sided z-derivative from rows 1-4
          sx = A * (FA(k.TXX1, i + 1, 0) - FA(k.TXX1, i - 2, 0)) + B * (FA(k.TXX1, i, 0) - FA(k.TXX1, i - 1, 0)) +
               c17_24 * FA(k.TXZ1, i, 1) + c3_8 * FA(k.TXZ1, i, 2) - c5_24 * FA(k.TXZ1, i, 3) +
               c1_24 * FA(k.TXZ1, i, 4);
        } else {  // second-order build (inh_qx_int_0.F90:40-47)
          sx = (FA(k.TXX1, i, 0) - FA(k.TXX1, i - 1, 0)) + FA(k.TXZ1, i, 1);
        }
        const float px = (l == 0 && k.SECOND)
                             ? (FA(k.PR1, i, 0) - FA(k.PR1, i - 1, 0))
                             : A * (FA(k.PR1, i + 1, l) - FA(k.PR1, i - 2, l)) + B * (FA(k.PR1, i, l) - FA(k.PR1, i - 1, l));
        qn = v0.x * q - v0.y * sx - v0.z * px;
        un = u + v0.w * q + v1.x * sx + v1.y * px;
      } else {
        const bool right = i > 3;
        const int j = right ? i - (MX - 1) : 3 - i;
        const int t0 = right ? 4 : 0;
        const float oa = XL2(t0, j, l), ob = XL2(t0 + 1, j, l), oc = XL2(t0 + 2, j, l), od = XL2(t0 + 3, j, l);
        const size_t mo = right ? (size_t)l * (T + 2) + (i - (MX - 1)) : (size_t)l * (T + 3) + ix;
        float *G = right ? k.GXXH : k.GXXL, *H = right ? k.HXXH : k.HXXL;
        float *R = right ? k.RXXH : k.RXXL, *S = right ? k.SXXH : k.SXXL;
        const float g = G[mo], hh = H[mo], r = R[mo], s = S[mo];
        const float dxx = FA(k.TXX1, i, l) - FA(k.TXX1, i - 1, l);
        const float pxx = FA(k.PR1, i, l) - FA(k.PR1, i - 1, l);
        const float tz1 = FA(k.TXZ1, i, l + 1), tz0 = FA(k.TXZ1, i, l);
        if (l != 0) {
          qn = v0.x * q - v0.y * (oa * (od * dxx + g) + tz1 - tz0) - v0.z * (oa * (od * pxx + hh));
          un = u + v0.w * q + v1.x * (oa * (od * dxx + r) + tz1 - tz0) + v1.y * (oa * (od * pxx + s));
        } else {  // row 0: the z term is TXZ(i,1) alone (inh_qx_2nd_0.F90, inh_v_u_2nd_0.F90)
          qn = v0.x * q - v0.y * (oa * (od * dxx + g) + tz1) - v0.z * (oa * (od * pxx + hh));
          un = u + v0.w * q + v1.x * (oa * (od * dxx + r) + tz1) + v1.y * (oa * (od * pxx + s));
        }
        if (ST) {
          G[mo] = ob * g + oc * dxx;
          H[mo] = ob * hh + oc * pxx;
          R[mo] = ob * r + oc * dxx;
          S[mo] = ob * s + oc * pxx;
        }
      }
    } else {
      // INH_QX_INT_2ND_PML (inh_qx_int_2nd_pml.F90:60-143), INH_V_U_INT_2ND_PML
      // (inh_v_u_int_2nd_pml.F90:62-143)
      const float dxx = FA(k.TXX1, i, l) - FA(k.TXX1, i - 1, l);
      const float dxz = FA(k.TXZ1, i, l + 1) - FA(k.TXZ1, i, l);
      const float pxx = FA(k.PR1, i, l) - FA(k.PR1, i - 1, l);
      const float za = ZH1(0, LL, ix), zb = ZH1(1, LL, ix), zc = ZH1(2, LL, ix), zd = ZH1(3, LL, ix);
      const size_t zo = (size_t)LL * NX + ix;
      const float gz = k.GXZH[zo], rz = k.RXZH[zo];
      if (inner) {
        qn = v0.x * q - v0.y * (za * (zd * dxz + gz) + dxx) - v0.z * pxx;
        un = u + v0.w * q + v1.x * (dxx + za * (zd * dxz + rz)) + v1.y * pxx;
      } else {
        const bool right = i > 3;
        const int j = right ? i - (MX - 1) : 3 - i;
        const int t0 = right ? 4 : 0;
        const float oa = XL2(t0, j, l), ob = XL2(t0 + 1, j, l), oc = XL2(t0 + 2, j, l), od = XL2(t0 + 3, j, l);
        const size_t mo = right ? (size_t)l * (T + 2) + (i - (MX - 1)) : (size_t)l * (T + 3) + ix;
        float *G = right ? k.GXXH : k.GXXL, *H = right ? k.HXXH : k.HXXL;
        float *R = right ? k.RXXH : k.RXXL, *S = right ? k.SXXH : k.SXXL;
        const float g = G[mo], hh = H[mo], r = R[mo], s = S[mo];
        qn = v0.x * q - v0.y * (oa * (od * dxx + g)) - v0.y * (za * (zd * dxz + gz)) - v0.z * (oa * (od * pxx + hh));
        un = u + v0.w * q + v1.x * (oa * (od * dxx + r) + za * (zd * dxz + rz)) + v1.y * (oa * (od * pxx + s));
        if (ST) {
          G[mo] = ob * g + oc * dxx;
          H[mo] = ob * hh + oc * pxx;
          R[mo] = ob * r + oc * dxx;
          S[mo] = ob * s + oc * pxx;
        }
      }
      if (ST) {
        k.GXZH[zo] = zb * gz + zc * dxz;
        k.RXZH[zo] = zb * rz + zc * dxz;
      }
    }
    o[0] = qn;
    o[1] = un;
  } else if (OUT) {
    o[0] = FA(k.QX1, i, l);
    o[1] = FA(k.U1, i, l);
  }

  // ---- QZ and W, x in [2-T, MX-1+T]
  if (i >= 2 - T && i <= MX - 1 + T) {
    const float q = FA(k.QZ0, i, l);
    const float w = FA(k.W0, i, l);
    float qn, wn;
    const bool inner = (i >= 3 && i <= MX - 2);
    if (!pmlrow) {
      if (inner) {
        // INH_QZ_INT (inh_qz_int.F90:41-55), INH_V_W_INT (inh_v_w_int.F90:43-58)
        float sz, pz;
        if (l >= 2) {
          sz = A * (FA(k.TZZ1, i, l + 1) - FA(k.TZZ1, i, l - 2) + FA(k.TXZ1, i + 2, l) - FA(k.TXZ1, i - 1, l)) +
               B * (FA(k.TZZ1, i, l) - FA(k.TZZ1, i, l - 1) + FA(k.TXZ1, i + 1, l) - FA(k.TXZ1, i, l));
          pz = A * (FA(k.PR1, i, l + 1) - FA(k.PR1, i, l - 2)) + B * (FA(k.PR1, i, l) - FA(k.PR1, i, l - 1));
        } else if (l == 0) {  // INH_QZ_INT_0, INH_V_W_INT_0: TZZ, PRES vanish at the surface
          sz = c35_8 * FA(k.TZZ1, i, 0) - c35_24 * FA(k.TZZ1, i, 1) + c21_40 * FA(k.TZZ1, i, 2) -
               c5_56 * FA(k.TZZ1, i, 3);
          pz = c35_8 * FA(k.PR1, i, 0) - c35_24 * FA(k.PR1, i, 1) + c21_40 * FA(k.PR1, i, 2) -
               c5_56 * FA(k.PR1, i, 3);
        } else {  // INH_V_QZ_INT_1, INH_V_W_INT_1
          sz = A * (FA(k.TXZ1, i + 2, 1) - FA(k.TXZ1, i - 1, 1)) + B * (FA(k.TXZ1, i + 1, 1) - FA(k.TXZ1, i, 1)) -
               c31_24 * FA(k.TZZ1, i, 0) + c29_24 * FA(k.TZZ1, i, 1) - c3_40 * FA(k.TZZ1, i, 2) +
               c1_168 * FA(k.TZZ1, i, 3);
          pz = -(c31_24 * FA(k.PR1, i, 0)) + c29_24 * FA(k.PR1, i, 1) - c3_40 * FA(k.PR1, i, 2) +
               c1_168 * FA(k.PR1, i, 3);
        }
        qn = v1.z * q - v1.w * sz - v2.x * pz;
        wn = w + v2.y * q + v2.z * sz + v2.w * pz;
      } else if (l == 0) {
        // INH_QZ_2ND_0, INH_V_W_2ND_0 (inh_qz_2nd_0.F90:48-80, inh_v_w_2nd_0.F90:50-84):
        // no x term on the surface row, so no PML memory
        const float sz = 3.0f * FA(k.TZZ1, i, 0) - c1_3 * FA(k.TZZ1, i, 1);
        const float pz = 3.0f * FA(k.PR1, i, 0) - c1_3 * FA(k.PR1, i, 1);
        qn = v1.z * q - v1.w * sz - v2.x * pz;
        wn = w + v2.y * q + v2.z * sz + v2.w * pz;
      } else {
        // INH_QZ_2ND (inh_qz_2nd.F90:49-88), INH_V_W_2ND (inh_v_w_2nd.F90:52-91)
        const bool right = i > 2;
        const int j = right ? i - (MX - 1) : 2 - i;
        const int t0 = right ? 4 : 0;
        const float oa = XL1(t0, j, l), ob = XL1(t0 + 1, j, l), oc = XL1(t0 + 2, j, l), od = XL1(t0 + 3, j, l);
        const size_t mo = right ? (size_t)l * (T + 2) + (i - (MX - 1)) : (size_t)l * (T + 2) + ix;
        float *G = right ? k.GZXH : k.GZXL, *R = right ? k.RZXH : k.RZXL;
        const float g = G[mo], r = R[mo];
        const float dxz = FA(k.TXZ1, i + 1, l) - FA(k.TXZ1, i, l);
        const float dzz = FA(k.TZZ1, i, l) - FA(k.TZZ1, i, l - 1);
        const float pzz = FA(k.PR1, i, l) - FA(k.PR1, i, l - 1);
        qn = v1.z * q - v1.w * (dzz + oa * (od * dxz + g)) - v2.x * pzz;
        wn = w + v2.y * q + v2.z * (dzz + oa * (od * dxz + r)) + v2.w * pzz;
        if (ST) {
          G[mo] = ob * g + oc * dxz;
          R[mo] = ob * r + oc * dxz;
        }
      }
    } else {
      // INH_QZ_INT_2ND_PML (inh_qz_int_2nd_pml.F90:57-144), INH_V_W_INT_2ND_PML
      // (inh_v_w_int_2nd_pml.F90:61-146)
      const float dxz = FA(k.TXZ1, i + 1, l) - FA(k.TXZ1, i, l);
      const float dzz = FA(k.TZZ1, i, l) - FA(k.TZZ1, i, l - 1);
      const float pzz = FA(k.PR1, i, l) - FA(k.PR1, i, l - 1);
      const float za = ZH2(0, LL, ix), zb = ZH2(1, LL, ix), zc = ZH2(2, LL, ix), zd = ZH2(3, LL, ix);
      const size_t zo = (size_t)LL * NX + ix;
      const float gz = k.GZZH[zo], hz = k.HZZH[zo], rz = k.RZZH[zo], sz = k.SZZH[zo];
      if (inner) {
        qn = v1.z * q - v1.w * (za * (zd * dzz + gz) + dxz) - v2.x * (za * (zd * pzz + hz));
        wn = w + v2.y * q + v2.z * (za * (zd * dzz + rz) + dxz) + v2.w * (za * (zd * pzz + sz));
      } else {
        const bool right = i > 2;
        const int j = right ? i - (MX - 1) : 2 - i;
        const int t0 = right ? 4 : 0;
        const float oa = XL1(t0, j, l), ob = XL1(t0 + 1, j, l), oc = XL1(t0 + 2, j, l), od = XL1(t0 + 3, j, l);
        const size_t mo = right ? (size_t)l * (T + 2) + (i - (MX - 1)) : (size_t)l * (T + 2) + ix;
        float *G = right ? k.GZXH : k.GZXL, *R = right ? k.RZXH : k.RZXL;
        const float g = G[mo], r = R[mo];
        qn = v1.z * q - v1.w * (za * (zd * dzz + gz)) - v1.w * (oa * (od * dxz + g)) - v2.x * (za * (zd * pzz + hz));
        wn = w + v2.y * q + v2.z * (za * (zd * dzz + rz) + oa * (od * dxz + r)) + v2.w * (za * (zd * pzz + sz));
        if (ST) {
          G[mo] = ob * g + oc * dxz;
          R[mo] = ob * r + oc * dxz;
        }
      }
      if (ST) {
        k.GZZH[zo] = zb * gz + zc * dzz;
        k.HZZH[zo] = zb * hz + zc * pzz;
        k.RZZH[zo] = zb * rz + zc * dzz;
        k.SZZH[zo] = zb * sz + zc * pzz;
      }
    }
    o[2] = qn;
    o[3] = wn;
  } else if (OUT) {
    o[2] = FA(k.QZ1, i, l);
    o[3] = FA(k.W1, i, l);
  }
}

// The velocity half-step of cell (i,l): the new values stored, the PML
// memories updated.
__device__ __forceinline__ void velocity_cell(const KParams &k, const int i, const int l) {
  float o[4];
  velocity_cell_t<true, false>(k, i, l, o);
  if (i >= 3 - k.T && i <= k.MX - 1 + k.T) {
    FA(k.QX1, i, l) = o[0];
    FA(k.U1, i, l) = o[1];
  }
  if (i >= 2 - k.T && i <= k.MX - 1 + k.T) {
    FA(k.QZ1, i, l) = o[2];
    FA(k.W1, i, l) = o[3];
  }
}


// ------------------------------------------------------- interior fast path
// Interior cells (4 <= i <= MX-2, rows of the IC class away from the image
// rows) use only the 4th-order formulas. A wavefront owns 64 columns of a run
// of rows; the z-stencil values of its own column are loaded once per run and
// kept in registers, and the x-neighbours come from the adjacent lanes through
// cross-lane shuffles, with the 2 columns on either side of the 64 loaded by
// four halo lanes. The arithmetic is the same expression as the per-cell path.

// Values of a row at columns i-2, i-1, i+1, i+2 for lane i of a 64-lane tile
// starting at column c; h holds the halo: lane 0 -> c+64, 1 -> c+65, 2 -> c-1, 3 -> c-2.
struct XNb {
  float m2, m1, p1, p2;
};
__device__ __forceinline__ XNb xneighbours(float v, float h, int lane) {
  const float hr0 = __shfl(h, 0), hr1 = __shfl(h, 1), hl1 = __shfl(h, 2), hl2 = __shfl(h, 3);
  const float sp1 = __shfl(v, (lane + 1) & 63), sp2 = __shfl(v, (lane + 2) & 63);
  const float sm1 = __shfl(v, (lane + 63) & 63), sm2 = __shfl(v, (lane + 62) & 63);
  XNb r;
  r.p1 = lane == 63 ? hr0 : sp1;
  r.p2 = lane == 63 ? hr1 : (lane == 62 ? hr0 : sp2);
  r.m1 = lane == 0 ? hl1 : sm1;
  r.m2 = lane == 0 ? hl2 : (lane == 1 ? hl1 : sm2);
  return r;
}

// Short unrolled column runs: a block of 4 wavefronts covers 64 columns x
// 4*NR rows, wavefront w taking rows l0+w*NR .. l0+w*NR+NR-1 of its lanes'
// columns. Loads of one run are independent of each other, so all of them can
// be in flight at once; the z-stencil rows shared by the NR cells of a run are
// loaded once. Blocks are numbered so that each XCD (block b runs on XCD b%8)
// gets a contiguous band of tiles, keeping the halo rows shared by vertically
// adjacent tiles inside one L2.
__device__ __forceinline__ int xcd_tile(int b, int nblocks) {
  const int per = nblocks >> 3, rem = nblocks & 7, x = b & 7, q = b >> 3;
  // XCD x owns tiles [start_x, start_x + per + (x < rem)) in row-major order
  return x * per + min(x, rem) + q;
}

// A run's cells [x0, xl) x [z0, z1) in ntx strips of 64 columns; lanes at
// columns >= xl (the last strip of a width that is not a multiple of 64)
// compute but do not store, and their loads are clamped to the grid's last
// column (the lanes below xl never use those values).
template <int NR>
__device__ __forceinline__ void stress_run(const KParams &k, int x0, int xl, int ntx, int z0,
                                           int z1, int tile, int wv, int lane, int itile) {
  const int tx = tile % ntx, tz = tile / ntx;
  const int c = x0 + tx * 64;
  const int i = c + lane;
  const int l0 = z0 + (tz * 4 + wv) * NR;
  if (l0 >= z1) return;
  const bool st = i < xl;
  const int P = k.P;
  const int ix = min(i + k.OFF, k.NX - 1);
  const int hx = min((lane == 0 ? c + 64 : lane == 1 ? c + 65 : lane == 2 ? c - 1 : lane == 3 ? c - 2 : i) + k.OFF,
                     k.NX - 1);
  const float *__restrict__ U = k.U0;
  const float *__restrict__ W = k.W0;
  const float *__restrict__ QX = k.QX0;
  const float *__restrict__ QZ = k.QZ0;
  const float *__restrict__ TXX = k.TXX0;
  const float *__restrict__ TZZ = k.TZZ0;
  const float *__restrict__ TXZ = k.TXZ0;
  const float *__restrict__ PR = k.PR0;
  float *__restrict__ nTXX = k.TXX1;
  float *__restrict__ nTZZ = k.TZZ1;
  float *__restrict__ nTXZ = k.TXZ1;
  float *__restrict__ nPR = k.PR1;
  const float AH = k.AH, BH = k.BH, BC = k.BC, C = k.CH, DT = k.DT;
  // rows l0-2 .. l0+NR+1 of U, W, QZ at this column (z-stencils of the run)
  // (rows past the last row of a partial run are clamped into the grid; their
  // values are never used)
  float u[NR + 3], w[NR + 3], z[NR + 3];
#pragma unroll
  for (int r = 0; r < NR + 3; ++r) {
    const int ru = min(l0 - 2 + r, k.PMLZH), rw = min(l0 - 1 + r, k.PMLZH);
    u[r] = U[ru * P + ix];
    w[r] = W[rw * P + ix];   // w[r] = W(l0-1+r)
    z[r] = QZ[rw * P + ix];
  }
#pragma unroll
  for (int j = 0; j < NR; ++j) {
    const int l = l0 + j;
    if (l >= z1) break;
    const int r = l * P;
    const float qx0 = QX[r + ix];
    const XNb un = xneighbours(u[j + 2], U[r + hx], lane);
    const XNb wn = xneighbours(w[j + 1], W[r + hx], lane);
    const XNb qn = xneighbours(qx0, QX[r + hx], lane);
    const int jm = jm_at(k, r + ix);
    const float4 s0 = k.stab[3 * jm], s1 = k.stab[3 * jm + 1], s2 = k.stab[3 * jm + 2];
    const float um2 = u[j], um1 = u[j + 1], u0 = u[j + 2], up1 = u[j + 3];
    const float wm1 = w[j], w0 = w[j + 1], wp1 = w[j + 2], wp2 = w[j + 3];
    const float zm1 = z[j], z0 = z[j + 1], zp1 = z[j + 2], zp2 = z[j + 3];
    const float exz = AH * (up1 - um2 + wn.p1 - wn.m2) + BH * (u0 - um1 + w0 - wn.m1);
    const float exx = AH * (un.p2 - un.m1) + BH * (un.p1 - u0);
    const float qxx = AH * (qn.p2 - qn.m1) + BH * (qn.p1 - qx0);
    float ezz = (C * (wp1 - w0));
    float qzz = (C * (zp1 - z0));
    ezz = AH * (wp2 - wm1) + BC * ezz;
    qzz = AH * (zp2 - zm1) + BC * qzz;
    if (!st) continue;
    nTXZ[r + ix] = TXZ[r + ix] + DT * (s0.x * exz);
    const float rxx = s0.y * exx + s0.z * ezz + s0.w * qxx + s0.w * qzz;
    const float rzz = s1.x * exx + s1.y * ezz + s1.z * qxx + s1.z * qzz;
    const float rpp = s1.w * exx + s2.x * ezz + s2.y * qxx + s2.y * qzz;
    if (i == k.LIN_X && l == k.LIN_Z) {
      const float2 ds = k.dsrc[itile - 1];
      nTXX[r + ix] = TXX[r + ix] + DT * (rxx - ds.x);
      nTZZ[r + ix] = TZZ[r + ix] + DT * (rzz - ds.x);
      nPR[r + ix] = PR[r + ix] + DT * (rpp + ds.y);
    } else {
      nTXX[r + ix] = TXX[r + ix] + DT * rxx;
      nTZZ[r + ix] = TZZ[r + ix] + DT * rzz;
      nPR[r + ix] = PR[r + ix] + DT * rpp;
    }
  }
}

template <int NR>
__device__ __forceinline__ void velocity_run(const KParams &k, int x0, int xl, int ntx, int z0,
                                             int z1, int tile, int wv, int lane) {
  const int tx = tile % ntx, tz = tile / ntx;
  const int c = x0 + tx * 64;
  const int i = c + lane;
  const int l0 = z0 + (tz * 4 + wv) * NR;
  if (l0 >= z1) return;
  const bool st = i < xl;
  const int P = k.P;
  const int ix = min(i + k.OFF, k.NX - 1);
  const int hx = min((lane == 0 ? c + 64 : lane == 1 ? c + 65 : lane == 2 ? c - 1 : lane == 3 ? c - 2 : i) + k.OFF,
                     k.NX - 1);
  const float *__restrict__ TXX = k.TXX1;
  const float *__restrict__ TZZ = k.TZZ1;
  const float *__restrict__ TXZ = k.TXZ1;
  const float *__restrict__ PR = k.PR1;
  const float *__restrict__ U = k.U0;
  const float *__restrict__ W = k.W0;
  const float *__restrict__ QX = k.QX0;
  const float *__restrict__ QZ = k.QZ0;
  float *__restrict__ nU = k.U1;
  float *__restrict__ nW = k.W1;
  float *__restrict__ nQX = k.QX1;
  float *__restrict__ nQZ = k.QZ1;
  const float A = k.A, B = k.B;
  // rows l0-2 .. l0+NR+1: xz[r] = TXZ(l0-2+r), zz[r] = TZZ(l0-2+r), pp[r] = PRES(l0-2+r)
  // (rows past the last row of a partial run are clamped into the grid; their
  // values are never used)
  float xz[NR + 4], zz[NR + 3], pp[NR + 3];
#pragma unroll
  for (int r = 0; r < NR + 4; ++r) xz[r] = (r == 0) ? 0.0f : TXZ[min(l0 - 2 + r, k.PMLZH) * P + ix];
#pragma unroll
  for (int r = 0; r < NR + 3; ++r) {
    const int rr = min(l0 - 2 + r, k.PMLZH);
    zz[r] = TZZ[rr * P + ix];
    pp[r] = PR[rr * P + ix];
  }
#pragma unroll
  for (int j = 0; j < NR; ++j) {
    const int l = l0 + j;
    if (l >= z1) break;
    const int r = l * P;
    const float txx0 = TXX[r + ix];
    const float xz0 = xz[j + 2], p0 = pp[j + 2];
    const XNb xx = xneighbours(txx0, TXX[r + hx], lane);
    const XNb xzn = xneighbours(xz0, TXZ[r + hx], lane);
    const XNb pn = xneighbours(p0, PR[r + hx], lane);
    const int jm = jm_at(k, r + ix);
    const float4 v0 = k.vtab[3 * jm], v1 = k.vtab[3 * jm + 1], v2 = k.vtab[3 * jm + 2];
    if (!st) continue;
    const float q = QX[r + ix], u = U[r + ix], qz = QZ[r + ix], w = W[r + ix];
    const float sx = A * (xx.p1 - xx.m2 + xz[j + 4] - xz[j + 1]) + B * (txx0 - xx.m1 + xz[j + 3] - xz0);
    const float px = A * (pn.p1 - pn.m2) + B * (p0 - pn.m1);
    nQX[r + ix] = v0.x * q - v0.y * sx - v0.z * px;
    nU[r + ix] = u + v0.w * q + v1.x * sx + v1.y * px;
    const float sz = A * (zz[j + 3] - zz[j] + xzn.p2 - xzn.m1) + B * (zz[j + 2] - zz[j + 1] + xzn.p1 - xz0);
    const float pz = A * (pp[j + 3] - pp[j]) + B * (p0 - pp[j + 1]);
    nQZ[r + ix] = v1.z * qz - v1.w * sz - v2.x * pz;
    nW[r + ix] = w + v2.y * qz + v2.z * sz + v2.w * pz;
  }
}

constexpr int kBlock = 256;  // 4 wavefronts
constexpr int kRun = 4;      // rows per wavefront run
constexpr int kRunCells = 2; // rows per wavefront run of the partition without fused tiles
constexpr int kFR = 128;     // rows per fused block (the tallest; bq_create picks per grid)

__device__ __forceinline__ bool rest_cell(const Split &sp, int g, int *i, int *l) {
  if (g >= sp.cum[kMaxCells]) return false;
  int r = 0;
  while (g >= sp.cum[r + 1]) ++r;
  const int off = g - sp.cum[r];
  const int w = sp.rx1[r] - sp.rx0[r];
  *i = sp.rx0[r] + off % w;
  *l = sp.rz0[r] + off / w;
  return true;
}

// Block b of the interior runs -> rectangle and tile of that rectangle.
__device__ __forceinline__ int run_rect(const Split &sp, int b, int *tile) {
  int r = 0;
  while (r + 1 < sp.nrect && b >= sp.qcum[r + 1]) ++r;
  *tile = b - sp.qcum[r];
  return r;
}

// -------------------------------------------------------------- fused step
// Interior cells far enough from every boundary that both half-steps use only
// their 4th-order formulas are advanced by one kernel that does the stress
// and the velocity half-step together. A block owns kFW columns x NR rows and
// marches down them: at row s it forms the new stresses of row s (its own
// columns and two columns on either side), then the new velocities of row
// s-2, whose stencils reach new stresses two rows and two columns away. The
// stresses of the two rows above and below the block and of the side columns
// are recomputed from the current fields, not stored; the rows a block owns
// are written once. Old velocity rows and new stress rows pass through LDS
// rings; the block's own column keeps its z-windows in registers. Fields are
// read from the current set and written to the next, so one step moves
// 8 fields in, 8 fields out and one material index per cell.
constexpr int kFW = 512;       // columns per block, one thread each
constexpr int kLW = kFW + 8;   // old-velocity LDS row: columns c-4 .. c+kFW+3
constexpr int kSW = kFW + 4;   // new-stress LDS row: columns c-2 .. c+kFW+1
constexpr int kStepBlock = kFW;  // threads of a k_step block

// New TXX, TZZ, TXZ, PRES at LDS column q (global column i) of row s, from
// rings of old-velocity rows (row r of U in slot r % DU, of W in r % DW, of
// QZ in r % DZ, of QX in r % DQ; LW floats per row) - the expressions of
// stress_run. jm: the cell's material; itile: the step (source term).
template <int DU, int DW, int DZ, int DQ, int LW>
__device__ __forceinline__ void ring_stress_t(const KParams &k, const float (*sU)[LW],
                                              const float (*sW)[LW], const float (*sZ)[LW],
                                              const float (*sQ)[LW], int s, int q, int i,
                                              float txx, float tzz, float txz, float pr,
                                              const float4 s0, const float4 s1, const float4 s2,
                                              int itile, float *o) {
  const int u0s = s % DU, w0s = s % DW, z0s = s % DZ, qs = s % DQ;
  const float um2 = sU[(s - 2) % DU][q], um1 = sU[(s - 1) % DU][q], u0 = sU[u0s][q];
  const float up1 = sU[(s + 1) % DU][q];
  const float uxm1 = sU[u0s][q - 1], uxp1 = sU[u0s][q + 1], uxp2 = sU[u0s][q + 2];
  const float wm1 = sW[(s - 1) % DW][q], w0 = sW[w0s][q], wp1 = sW[(s + 1) % DW][q];
  const float wp2 = sW[(s + 2) % DW][q];
  const float wxm2 = sW[w0s][q - 2], wxm1 = sW[w0s][q - 1], wxp1 = sW[w0s][q + 1];
  const float zm1 = sZ[(s - 1) % DZ][q], z0 = sZ[z0s][q], zp1 = sZ[(s + 1) % DZ][q];
  const float zp2 = sZ[(s + 2) % DZ][q];
  const float qxm1 = sQ[qs][q - 1], qx0 = sQ[qs][q], qxp1 = sQ[qs][q + 1], qxp2 = sQ[qs][q + 2];
  const float AH = k.AH, BH = k.BH, BC = k.BC, C = k.CH, DT = k.DT;
  const float exz = AH * (up1 - um2 + wxp1 - wxm2) + BH * (u0 - um1 + w0 - wxm1);
  const float exx = AH * (uxp2 - uxm1) + BH * (uxp1 - u0);
  const float qxx = AH * (qxp2 - qxm1) + BH * (qxp1 - qx0);
  float ezz = (C * (wp1 - w0));
  float qzz = (C * (zp1 - z0));
  ezz = AH * (wp2 - wm1) + BC * ezz;
  qzz = AH * (zp2 - zm1) + BC * qzz;
  const float rxx = s0.y * exx + s0.z * ezz + s0.w * qxx + s0.w * qzz;
  const float rzz = s1.x * exx + s1.y * ezz + s1.z * qxx + s1.z * qzz;
  const float rpp = s1.w * exx + s2.x * ezz + s2.y * qxx + s2.y * qzz;
  o[2] = txz + DT * (s0.x * exz);
  if (i == k.LIN_X && s == k.LIN_Z) {
    const float2 ds = k.dsrc[itile - 1];
    o[0] = txx + DT * (rxx - ds.x);
    o[1] = tzz + DT * (rzz - ds.x);
    o[3] = pr + DT * (rpp + ds.y);
  } else {
    o[0] = txx + DT * rxx;
    o[1] = tzz + DT * rzz;
    o[3] = pr + DT * rpp;
  }
}

// ring_stress_t with the material's stress table entries read here.
template <int DU, int DW, int DZ, int DQ, int LW>
__device__ __forceinline__ void ring_stress(const KParams &k, const float (*sU)[LW],
                                            const float (*sW)[LW], const float (*sZ)[LW],
                                            const float (*sQ)[LW], int s, int q, int i, float txx,
                                            float tzz, float txz, float pr, int jm, int itile,
                                            float *o) {
  const float4 s0 = k.stab[3 * jm], s1 = k.stab[3 * jm + 1], s2 = k.stab[3 * jm + 2];
  ring_stress_t<DU, DW, DZ, DQ, LW>(k, sU, sW, sZ, sQ, s, q, i, txx, tzz, txz, pr, s0, s1, s2,
                                    itile, o);
}

// New QX, U, QZ, W of one interior cell (INH_QX_INT, INH_V_UV_INT, INH_QZ_INT,
// INH_V_W_INT: the expressions of velocity_run). X, XZ, PP: the new TXX, TXZ,
// PRES of the cell's row at columns i-2 .. i+2 (element 2 is column i);
// xz[0..3]: new TXZ of rows l-1 .. l+2, zz[0..3] and pp[0..3]: new TZZ and
// PRES of rows l-2 .. l+1 at column i; q, u, qz, w: the old velocities.
// o = {QX, U, QZ, W}.
__device__ __forceinline__ void ring_velocity_t(const KParams &k, const float *X, const float *XZ,
                                                const float *PP, float xz0, float xz1, float xz2,
                                                float xz3, float zz0, float zz1, float zz2,
                                                float zz3, float pp0, float pp1, float pp2,
                                                float pp3, float q, float u, float qz, float w,
                                                const float4 v0, const float4 v1, const float4 v2,
                                                float *o) {
  const float A = k.A, B = k.B;
  const float xm2 = X[0], xm1 = X[1], x0 = X[2], xp1 = X[3];
  const float xzm1 = XZ[1], xzc = XZ[2], xzp1 = XZ[3], xzp2 = XZ[4];
  const float pm2 = PP[0], pm1 = PP[1], p0 = PP[2], pp1x = PP[3];
  // TXZ(l-1..l+2) = xz0..xz3, TZZ and PRES (l-2..l+1) = zz0..zz3, pp0..pp3 at column i
  const float sx = A * (xp1 - xm2 + xz3 - xz0) + B * (x0 - xm1 + xz2 - xz1);
  const float px = A * (pp1x - pm2) + B * (p0 - pm1);
  o[0] = v0.x * q - v0.y * sx - v0.z * px;
  o[1] = u + v0.w * q + v1.x * sx + v1.y * px;
  const float sz = A * (zz3 - zz0 + xzp2 - xzm1) + B * (zz2 - zz1 + xzp1 - xzc);
  const float pz = A * (pp3 - pp0) + B * (pp2 - pp1);
  o[2] = v1.z * qz - v1.w * sz - v2.x * pz;
  o[3] = w + v2.y * qz + v2.z * sz + v2.w * pz;
}

// ring_velocity_t with the material's velocity table entries read here.
__device__ __forceinline__ void ring_velocity(const KParams &k, const float *X, const float *XZ,
                                              const float *PP, float xz0, float xz1, float xz2,
                                              float xz3, float zz0, float zz1, float zz2, float zz3,
                                              float pp0, float pp1, float pp2, float pp3, float q,
                                              float u, float qz, float w, int jm, float *o) {
  const float4 v0 = k.vtab[3 * jm], v1 = k.vtab[3 * jm + 1], v2 = k.vtab[3 * jm + 2];
  ring_velocity_t(k, X, XZ, PP, xz0, xz1, xz2, xz3, zz0, zz1, zz2, zz3, pp0, pp1, pp2, pp3, q, u,
                  qz, w, v0, v1, v2, o);
}

// LDS of one march of W columns: old-velocity rings (U 5 rows, W and QZ 6,
// QX 4; columns c-4 .. c+W+3) and new-stress rings (TXX, TXZ, PRES 3 rows;
// columns c-2 .. c+W+1).
template <int W>
struct MarchLds {
  static constexpr int LW = W + 8, SW = W + 4;
  static constexpr int kFloats = 21 * LW + 9 * SW;
};

// A k_step block holds one full tile (kFW columns), two half tiles or four
// quarter tiles, each march with LDS of its own.
struct FusedLds {
  static constexpr int kMax1 = MarchLds<kFW>::kFloats, kMax2 = 2 * MarchLds<kFW / 2>::kFloats;
  static constexpr int kMax4 = 4 * MarchLds<kFW / 4>::kFloats;
  static constexpr int kFloats = kMax1 > kMax2 ? (kMax1 > kMax4 ? kMax1 : kMax4) : (kMax2 > kMax4 ? kMax2 : kMax4);
  float raw[kFloats];
};

// One march of the fused step over columns [c, c+W) and rows [a, e) by W
// threads (t = 0 .. W-1), with LDS at lds; only the first ws columns are
// stored (ws < W: a masked march, whose other threads form the halo values
// of the stored columns; their loads are clamped to the grid's last column).
// NS: rows of a march of full height. UNI: the march runs for NS rows
// whatever e is, doing no work past e, so that several marches of one block
// meet the same barriers.
template <int W, int NS, bool UNI>
__device__ __forceinline__ void fused_march(const KParams &k, int c, int ws, int a, int e, int itile,
                                            int t, float *lds) {
  constexpr int LW = MarchLds<W>::LW, SW = MarchLds<W>::SW;
  float (*sU)[LW] = reinterpret_cast<float (*)[LW]>(lds);
  float (*sW)[LW] = sU + 5;
  float (*sZ)[LW] = sW + 6;
  float (*sQ)[LW] = sZ + 6;
  float (*nX)[SW] = reinterpret_cast<float (*)[SW]>(lds + 21 * LW);
  float (*nXZ)[SW] = nX + 3;
  float (*nP)[SW] = nXZ + 3;
  const bool live = UNI ? a < e : true;         // an empty march only meets the barriers
  const int last = UNI ? a + NS + 1 : e + 1;    // the loop's last row
  const int P = k.P, OFF = k.OFF, GX = k.NX - 1;
  const int i = c + t;
  const bool own = t < ws;                       // stores its column
  const int gx = min(i + OFF, GX);
  const bool hl = t < 8;                         // loads one extra old-velocity column
  const int hq = t < 4 ? t : W + t;              // its LDS column (column c-4+hq)
  const int hgx = min(c - 4 + hq + OFF, GX);
  const bool hs = t < 4;                         // forms one extra column of new stresses
  const int hcol = t < 2 ? c - 2 + t : c + W + (t - 2);
  const int hgs = min(hcol + OFF, GX);
  const float *__restrict__ U = k.U0;
  const float *__restrict__ Wv = k.W0;
  const float *__restrict__ QX = k.QX0;
  const float *__restrict__ QZ = k.QZ0;
  // prefill the rings: U rows a-4..a-2, W and QZ rows a-3..a-1
  if (live)
    for (int r = a - 4; r <= a - 2; ++r) {
      sU[r % 5][t + 4] = U[r * P + gx];
      sW[(r + 1) % 6][t + 4] = Wv[(r + 1) * P + gx];
      sZ[(r + 1) % 6][t + 4] = QZ[(r + 1) * P + gx];
      if (hl) {
        sU[r % 5][hq] = U[r * P + hgx];
        sW[(r + 1) % 6][hq] = Wv[(r + 1) * P + hgx];
        sZ[(r + 1) % 6][hq] = QZ[(r + 1) * P + hgx];
      }
    }
  float wxz0 = 0, wxz1 = 0, wxz2 = 0, wxz3 = 0;             // new TXZ rows s-3..s
  float wzz0 = 0, wzz1 = 0, wzz2 = 0, wzz3 = 0, wzz4 = 0;   // new TZZ rows s-4..s
  float wp0 = 0, wp1 = 0, wp2 = 0, wp3 = 0, wp4 = 0;        // new PRES rows s-4..s
  int jm1 = 0, jm2 = 0, jm0 = 0;                            // JM rows s-2, s-1, s
  // values of row s, loaded one iteration ahead so that the loads of the next
  // row are in flight while this row is computed
  float pu = 0, pw = 0, pz = 0, pq = 0, fhu = 0, fhw = 0, fhz = 0, fhq = 0;
  float ftxx = 0, ftzz = 0, ftxz = 0, fpr = 0, fhtxx = 0, fhtzz = 0, fhtxz = 0, fhpr = 0;
  int fjm = 0, fhjm = 0;
  auto fetch = [&](int s) {
    const int r = s * P;
    pu = stream_load(&U[r + P + gx]); pw = stream_load(&Wv[r + 2 * P + gx]);
    pz = stream_load(&QZ[r + 2 * P + gx]); pq = stream_load(&QX[r + gx]);
    if (hl) { fhu = U[r + P + hgx]; fhw = Wv[r + 2 * P + hgx]; fhz = QZ[r + 2 * P + hgx]; fhq = QX[r + hgx]; }
    ftxx = stream_load(&k.TXX0[r + gx]); ftzz = stream_load(&k.TZZ0[r + gx]);
    ftxz = stream_load(&k.TXZ0[r + gx]); fpr = stream_load(&k.PR0[r + gx]);
    fjm = jm_at(k, r + gx);
    if (hs) {
      fhtxx = k.TXX0[r + hgs]; fhtzz = k.TZZ0[r + hgs]; fhtxz = k.TXZ0[r + hgs];
      fhpr = k.PR0[r + hgs]; fhjm = jm_at(k, r + hgs);
    }
  };
  if (live) fetch(a - 2);
  for (int s = a - 2; s <= last; ++s) {
    const bool act = UNI ? live && s <= e + 1 : true;   // this row has work
    const int r = s * P;
    float otxx = 0, otzz = 0, otxz = 0, opr = 0, htxx = 0, htzz = 0, htxz = 0, hpr = 0;
    int hjm = 0;
    if (act) {
      // ---- old velocities entering the rings, old stresses and index of row s
      sU[(s + 1) % 5][t + 4] = pu;
      sW[(s + 2) % 6][t + 4] = pw;
      sZ[(s + 2) % 6][t + 4] = pz;
      sQ[s & 3][t + 4] = pq;
      if (hl) {
        sU[(s + 1) % 5][hq] = fhu;
        sW[(s + 2) % 6][hq] = fhw;
        sZ[(s + 2) % 6][hq] = fhz;
        sQ[s & 3][hq] = fhq;
      }
      otxx = ftxx; otzz = ftzz; otxz = ftxz; opr = fpr;
      jm2 = jm1;
      jm1 = jm0;
      jm0 = fjm;
      htxx = fhtxx; htzz = fhtzz; htxz = fhtxz; hpr = fhpr;
      hjm = fhjm;
      if (s <= e) fetch(s + 1);
    }
    __syncthreads();
    // ---- new stresses of row s
    if (act) {
      float o[4];
      ring_stress<5, 6, 6, 4, LW>(k, sU, sW, sZ, sQ, s, t + 4, i, otxx, otzz, otxz, opr, jm0, itile, o);
      const int ns = s % 3;
      nX[ns][t + 2] = o[0];
      nXZ[ns][t + 2] = o[2];
      nP[ns][t + 2] = o[3];
      if (own && s >= a && s < e) {
        stream_store(&k.TXX1[r + gx], o[0]);
        stream_store(&k.TZZ1[r + gx], o[1]);
        stream_store(&k.TXZ1[r + gx], o[2]);
        stream_store(&k.PR1[r + gx], o[3]);
      }
      wxz0 = wxz1; wxz1 = wxz2; wxz2 = wxz3; wxz3 = o[2];
      wzz0 = wzz1; wzz1 = wzz2; wzz2 = wzz3; wzz3 = wzz4; wzz4 = o[1];
      wp0 = wp1; wp1 = wp2; wp2 = wp3; wp3 = wp4; wp4 = o[3];
      if (hs) {
        float ho[4];
        ring_stress<5, 6, 6, 4, LW>(k, sU, sW, sZ, sQ, s, hcol - (c - 4), hcol, htxx, htzz, htxz, hpr,
                                    hjm, itile, ho);
        const int hn = hcol - (c - 2);
        nX[ns][hn] = ho[0];
        nXZ[ns][hn] = ho[2];
        nP[ns][hn] = ho[3];
      }
    }
    __syncthreads();
    // ---- new velocities of row l = s-2 (INH_QX_INT, INH_V_UV_INT, INH_QZ_INT,
    //      INH_V_W_INT: the expressions of velocity_run)
    const int l = s - 2;
    if (act && own && l >= a && l < e) {
      const int sl = l % 3, rl = l * P;
      const float A = k.A, B = k.B;
      const float xm2 = nX[sl][t], xm1 = nX[sl][t + 1], x0 = nX[sl][t + 2], xp1 = nX[sl][t + 3];
      const float xzm1 = nXZ[sl][t + 1], xz0 = nXZ[sl][t + 2], xzp1 = nXZ[sl][t + 3];
      const float xzp2 = nXZ[sl][t + 4];
      const float pm2 = nP[sl][t], pm1 = nP[sl][t + 1], p0 = nP[sl][t + 2], pp1 = nP[sl][t + 3];
      const float q = sQ[l & 3][t + 4], u = sU[l % 5][t + 4];
      const float qz = sZ[l % 6][t + 4], w = sW[l % 6][t + 4];
      const float4 v0 = k.vtab[3 * jm2], v1 = k.vtab[3 * jm2 + 1], v2 = k.vtab[3 * jm2 + 2];
      // TXZ(l-1..l+2) = wxz0..wxz3, TZZ and PRES (l-2..l+1) = w*0..w*3 at column i
      const float sx = A * (xp1 - xm2 + wxz3 - wxz0) + B * (x0 - xm1 + wxz2 - wxz1);
      const float px = A * (pp1 - pm2) + B * (p0 - pm1);
      stream_store(&k.QX1[rl + gx], v0.x * q - v0.y * sx - v0.z * px);
      stream_store(&k.U1[rl + gx], u + v0.w * q + v1.x * sx + v1.y * px);
      const float sz = A * (wzz3 - wzz0 + xzp2 - xzm1) + B * (wzz2 - wzz1 + xzp1 - xz0);
      const float pz = A * (wp3 - wp0) + B * (wp2 - wp1);
      stream_store(&k.QZ1[rl + gx], v1.z * qz - v1.w * sz - v2.x * pz);
      stream_store(&k.W1[rl + gx], w + v2.y * qz + v2.z * sz + v2.w * pz);
    }
  }
}

// Fused block of launch block b (block b runs on XCD b % 8): XCD x takes a
// contiguous band of the full tiles, then of the half-tile blocks, then of
// the quarter-tile blocks, in that order, so the short blocks are dispatched
// on every XCD after its long ones and take the slots the half blocks free.
// -1: no block (past the XCD's share).
__device__ __forceinline__ int fused_item(const Fused &fu, int b) {
  const int x = b & 7;
  int q = b >> 3, base = 0;
  const int n[3] = {fu.nfx * fu.nfz, fu.nb2, fu.nb4};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int per = n[c] >> 3, rem = n[c] & 7;
    const int cnt = per + (x < rem);
    if (q < cnt) return base + x * per + min(x, rem) + q;
    q -= cnt;
    base += n[c];
  }
  return -1;
}

// Block b of the fused launch (kFW threads): a full tile (kFW columns, NR
// rows), or the two half tiles (kFW/2 columns, NR/2 rows) or four quarter
// tiles (kFW/4 columns, NR/4 rows) of one block of the narrow columns.
template <int NR>
__device__ __forceinline__ void fused_block(const KParams &k, const Fused &fu, int b, int itile,
                                            FusedLds &L) {
  const int nfull = fu.nfx * fu.nfz;
  if (b < nfull) {
    const int bx = b % fu.nfx, bz = b / fu.nfx;
    const int a = fu.vz0 + bz * NR;
    fused_march<kFW, NR, false>(k, fu.vx0 + bx * kFW, kFW, a, min(a + NR, fu.vz1), itile,
                                threadIdx.x, L.raw);
    return;
  }
  b -= nfull;
  const int c2 = fu.vx0 + kFW * fu.nfx;
  if (b < fu.nb2) {
    constexpr int W = kFW / 2, H = NR / 2;
    const int j = threadIdx.x / W;
    const int a = fu.vz0 + (2 * b + j) * H;
    fused_march<W, H, true>(k, c2, W, a, min(a + H, fu.vz1), itile, threadIdx.x % W,
                            L.raw + j * MarchLds<W>::kFloats);
    return;
  }
  b -= fu.nb2;
  constexpr int W = kFW / 4, H = NR / 4;
  const int per = fu.nb4 / fu.n4;    // blocks of one quarter column
  const int col = b / per;
  b -= col * per;
  const int j = threadIdx.x / W;
  const int a = fu.vz0 + (4 * b + j) * H;
  fused_march<W, H, true>(k, c2 + (kFW / 2) * fu.h2 + W * col, col == fu.n4 - 1 ? fu.w4 : W, a,
                          min(a + H, fu.vz1), itile, threadIdx.x % W,
                          L.raw + j * MarchLds<W>::kFloats);
}

// Receiver samples of the fields a step left behind (store_seismograms.F90:58-83,
// store_seismograms_pressure.F90:50-56); f points at the set holding them.
__device__ __forceinline__ void sample_receivers(const KParams &k, int j, int step, bool cur) {
  if (j >= k.MR) return;
  const int4 r = k.rec[j];
  const int nzl = k.PMLZH, lo = k.PMLL, hi = k.PMLL + k.NX - 1;
  const bool okx = r.x >= lo && r.x <= hi && r.y >= 0 && r.y <= nzl;
  const bool okz = r.z >= lo && r.z <= hi && r.w >= 0 && r.w <= nzl;
  const bool okp = r.z >= lo && r.z <= hi && r.y >= 0 && r.y <= nzl;
  const float *U = cur ? k.U0 : k.U1, *W = cur ? k.W0 : k.W1;
  const float *QX = cur ? k.QX0 : k.QX1, *QZ = cur ? k.QZ0 : k.QZ1;
  const float *PR = cur ? k.PR0 : k.PR1, *TXX = cur ? k.TXX0 : k.TXX1;
  float *o = k.traces + ((size_t)(step - 1) * k.MR + j) * 6;
  o[0] = okx ? FA(U, r.x, r.y) : 0.0f;
  o[1] = okz ? -FA(W, r.z, r.w) : 0.0f;
  o[2] = okx ? FA(QX, r.x, r.y) : 0.0f;
  o[3] = okz ? -FA(QZ, r.z, r.w) : 0.0f;
  o[4] = okp ? FA(PR, r.z, r.y) : 0.0f;
  o[5] = okp ? -FA(TXX, r.z, r.y) : 0.0f;
}

// ------------------------------------------------------------------- edges
// Row 1, stress imaging (HPLANE_1_VELOCITY, hplane_1_velocity.F90:61-76):
// INH_V_QX_INT_1_SI (inh_v_qx_int_1_si.F90:26-31) for all x first, then
// INH_W_QZ_INT_1_SI (inh_w_qz_int_1_si.F90:30-44), which reads the new row-1
// U and QX at i+1: each block recomputes the one halo column it needs.
// Row PMLZH (HPLANE_PMLZH0_VELOCITY, hplane_pmlzh0_velocity.F90:34-63) from
// the new rows PMLZH-1, PMLZH-2 and the two stored levels, then the history
// is advanced, as the store of time_loop.F90:133-155 sees it next step.
// Block bx of the edge rows; su, sq: kEdgeBlock + 1 floats of LDS each.
__device__ __forceinline__ void edges_block(const KParams &k, const int bx, float *su, float *sq) {
  const int T = k.T, MX = k.MX, NX = k.NX;
  const int ix0 = bx * kEdgeBlock;
  const int ix = ix0 + threadIdx.x;
  const int i = ix + k.PMLL;
  const bool col = ix < NX;
  const bool si = k.SI != 0;
  const int lst = k.PMLZH;
  // Every value this block reads is loaded first: rows 2, 3 (written by the
  // velocity half-step before this launch), rows PMLZH-2 .. PMLZH and the
  // history, none of which this launch writes (rows 1 and PMLZH >= 4 are its
  // only field stores), so the loads are in flight together.
  float* const F[4] = {k.U1, k.W1, k.QX1, k.QZ1};
  float bf0[4] = {0, 0, 0, 0}, bf1[4] = {0, 0, 0, 0}, bf2[4] = {0, 0, 0, 0};
  float bsc[4][3] = {}, bsp[4][3] = {};
  if (col) {
#pragma unroll
    for (int f = 0; f < 4; ++f) {
      const float *sc = k.SZc + ((size_t)f * NX + ix) * 3;
      const float *sp = k.SZp + ((size_t)f * NX + ix) * 3;
      bf0[f] = FA(F[f], i, lst);
      bf1[f] = FA(F[f], i, lst - 1);
      bf2[f] = FA(F[f], i, lst - 2);
      bsc[f][0] = sc[0]; bsc[f][1] = sc[1]; bsc[f][2] = sc[2];
      bsp[f][0] = sp[0]; bsp[f][1] = sp[1]; bsp[f][2] = sp[2];
    }
  }
  const bool p2 = si && col && i >= 3 - T && i <= MX - 2 + T;
  float up2 = 0, u2 = 0, w3 = 0, qz3 = 0, qxp2 = 0, qx2 = 0;
  float2 aux = make_float2(0.0f, 0.0f);
  if (p2) {
    up2 = FA(k.U1, i + 1, 2); u2 = FA(k.U1, i, 2);
    w3 = FA(k.W1, i, 3); qz3 = FA(k.QZ1, i, 3);
    qxp2 = FA(k.QX1, i + 1, 2); qx2 = FA(k.QX1, i, 2);
    aux = k.atab[jm_at(k, ix)];  // JM(I,0)
  }
  // phase 1: U(i,1), QX(i,1) for the block's columns and one halo column
  // (stress imaging only; without it row 1 is a per-cell row of k_velocity)
  for (int t = threadIdx.x; si && t <= kEdgeBlock; t += kEdgeBlock) {
    const int jx = ix0 + t;
    const int c = jx + k.PMLL;
    if (jx < NX && c >= 3 - T && c <= MX - 1 + T) {
      const float u1 = FA(k.U1, c, 2) + FA(k.W1, c, 2) - FA(k.W1, c - 1, 2);
      const float q1 = FA(k.QX1, c, 2) + FA(k.QZ1, c, 2) - FA(k.QZ1, c - 1, 2);
      su[t] = u1;
      sq[t] = q1;
    }
  }
  if (si) __syncthreads();
  if (col) {
    if (si && i >= 3 - T && i <= MX - 1 + T) {
      FA(k.U1, i, 1) = su[threadIdx.x];
      FA(k.QX1, i, 1) = sq[threadIdx.x];
    }
    // phase 2
    if (p2) {
      const float u0 = su[threadIdx.x], u1 = su[threadIdx.x + 1];
      const float q0 = sq[threadIdx.x], q1 = sq[threadIdx.x + 1];
      const float du = u1 - u0 + up2 - u2;
      FA(k.W1, i, 1) = w3 + aux.x * du;
      FA(k.QZ1, i, 1) = qz3 + aux.y * du + (q1 - q0 + qxp2 - qx2);
    }
    // bottom row
    const bool upd = i >= 1 - T && i <= MX + T;
#pragma unroll
    for (int f = 0; f < 4; ++f) {
      const float *c = (f == 1 || f == 3) ? k.abw : k.abh;
      float *sc = k.SZc + ((size_t)f * NX + ix) * 3;
      float *sp = k.SZp + ((size_t)f * NX + ix) * 3;
      const float f1 = bf1[f], f2 = bf2[f];
      sp[0] = bsc[f][0]; sp[1] = bsc[f][1]; sp[2] = bsc[f][2];
      if (upd) {
        const float n = c[0] * f1 + c[1] * f2 + c[2] * bsc[f][0] + c[3] * bsc[f][1] +
                        c[4] * bsc[f][2] + c[5] * bsp[f][0] + c[6] * bsp[f][1] + c[7] * bsp[f][2];
        FA(F[f], i, lst) = n;
        sc[0] = n; sc[1] = f1; sc[2] = f2;
      } else {
        // columns outside the bottom update keep their values; the history
        // still records them (time_loop.F90:140-153 copies whole rows)
        sc[0] = bf0[f]; sc[1] = f1; sc[2] = f2;
      }
    }
  }
}

// ------------------------------------------------ edge rows with the velocities
// The velocity half-step's edge blocks (Split::fold), kBlock = 2 * kEdgeCols
// threads: thread t takes column c + t % kEdgeCols of the block's first or
// second row (t / kEdgeCols).
static_assert(kBlock == 2 * kEdgeCols, "an edge block is two rows of kEdgeCols columns");
constexpr int kEdgeLds = 4 * (kEdgeCols + 2) + 2 * kEdgeCols + 2 * (kEdgeCols + 1);

// Top block bx: rows 2 and 3 of its columns [c, ce), then row 1 as
// HPLANE_1_VELOCITY forms it (hplane_1_velocity.F90:61-76: INH_V_QX_INT_1_SI,
// inh_v_qx_int_1_si.F90:26-31, for all x, then INH_W_QZ_INT_1_SI,
// inh_w_qz_int_1_si.F90:30-44, from the new rows 1-3). Row 1 at column i
// needs the new row 2 at i-1 .. i+1: the block recomputes row 2 at c-1 and
// ce (interior cells, no PML memory) without storing them.
__device__ __forceinline__ void top_block(const KParams &k, const Split &sp, int bx, float *lds) {
  constexpr int E = kEdgeCols;
  const int T = k.T, MX = k.MX, PX1 = k.PMLL + k.NX;
  const int last = sp.etop - 1;
  const int c = bx == last ? sp.elast : k.PMLL + E * bx;
  const int ce = bx == last ? PX1 : min(k.PMLL + E * (bx + 1), sp.elast);
  float *sQ2 = lds, *sU2 = lds + (E + 2), *sZ2 = lds + 2 * (E + 2), *sW2 = lds + 3 * (E + 2);
  float *sZ3 = lds + 4 * (E + 2), *sW3 = sZ3 + E, *su = sW3 + E, *sq = su + (E + 1);
  const int t = threadIdx.x, r = t / E, q = t % E;
  const int i = c + q;
  if (i < ce) {
    float o[4];
    velocity_cell_t<true, true>(k, i, 2 + r, o);
    if (i >= 3 - T && i <= MX - 1 + T) {
      FA(k.QX1, i, 2 + r) = o[0];
      FA(k.U1, i, 2 + r) = o[1];
    }
    if (i >= 2 - T && i <= MX - 1 + T) {
      FA(k.QZ1, i, 2 + r) = o[2];
      FA(k.W1, i, 2 + r) = o[3];
    }
    if (r == 0) {
      sQ2[q + 1] = o[0]; sU2[q + 1] = o[1]; sZ2[q + 1] = o[2]; sW2[q + 1] = o[3];
    } else {
      sZ3[q] = o[2]; sW3[q] = o[3];
    }
  }
  // row 2 of the neighbours' edge columns (second wave: its first two lanes)
  if (t == E || t == E + 1) {
    const int h = t == E ? c - 1 : ce;
    if (h >= k.PMLL && h < PX1) {
      float o[4];
      velocity_cell_t<false, true>(k, h, 2, o);
      const int qh = h - c + 1;
      sQ2[qh] = o[0]; sU2[qh] = o[1]; sZ2[qh] = o[2]; sW2[qh] = o[3];
    }
  }
  __syncthreads();
  // row 1, U and QX: columns c .. ce (ce for the next column's W and QZ)
  const int w = ce - c;
  if (t <= w) {
    const int j = c + t;
    if (j < PX1 && j >= 3 - T && j <= MX - 1 + T) {
      const float u1 = sU2[t + 1] + sW2[t + 1] - sW2[t];
      const float q1 = sQ2[t + 1] + sZ2[t + 1] - sZ2[t];
      su[t] = u1;
      sq[t] = q1;
      if (t < w) {
        FA(k.U1, j, 1) = u1;
        FA(k.QX1, j, 1) = q1;
      }
    }
  }
  __syncthreads();
  // row 1, W and QZ
  if (t < w && i >= 3 - T && i <= MX - 2 + T) {
    const int jm = jm_at(k, i + k.OFF);  // JM(I,0)
    const float2 aux = k.atab[jm];
    const float u0 = su[t], u1 = su[t + 1];
    const float q0 = sq[t], q1 = sq[t + 1];
    const float du = u1 - u0 + sU2[t + 2] - sU2[t + 1];
    FA(k.W1, i, 1) = sW3[t] + aux.x * du;
    FA(k.QZ1, i, 1) = sZ3[t] + aux.y * du + (q1 - q0 + sQ2[t + 2] - sQ2[t + 1]);
  }
}

// Bottom block bx: rows PMLZH-2 and PMLZH-1 of its columns, then row PMLZH
// (HPLANE_PMLZH0_VELOCITY, hplane_pmlzh0_velocity.F90:34-63) from them and the
// two stored levels, and the history advanced as the store of
// time_loop.F90:133-155 sees it next step.
__device__ __forceinline__ void bottom_block(const KParams &k, int bx, float *lds) {
  constexpr int E = kEdgeCols;
  const int T = k.T, MX = k.MX, NX = k.NX, lst = k.PMLZH;
  const int t = threadIdx.x, r = t / E, q = t % E;
  const int ix = bx * E + q;
  const int i = ix + k.PMLL;
  const bool col = ix < NX;
  float* const F[4] = {k.U1, k.W1, k.QX1, k.QZ1};
  if (col) {
    float o[4];
    velocity_cell_t<true, true>(k, i, lst - 2 + r, o);
    if (i >= 3 - T && i <= MX - 1 + T) {
      FA(k.QX1, i, lst - 2 + r) = o[0];
      FA(k.U1, i, lst - 2 + r) = o[1];
    }
    if (i >= 2 - T && i <= MX - 1 + T) {
      FA(k.QZ1, i, lst - 2 + r) = o[2];
      FA(k.W1, i, lst - 2 + r) = o[3];
    }
    // lds[(f * 2 + r) * E + q], fields in the order U, W, QX, QZ
    lds[(0 * 2 + r) * E + q] = o[1];
    lds[(1 * 2 + r) * E + q] = o[3];
    lds[(2 * 2 + r) * E + q] = o[0];
    lds[(3 * 2 + r) * E + q] = o[2];
  }
  __syncthreads();
  if (!col || r != 0) return;
  // the history and row PMLZH (nothing else writes them this step)
  const bool upd = i >= 1 - T && i <= MX + T;
#pragma unroll
  for (int f = 0; f < 4; ++f) {
    const float *cf = (f == 1 || f == 3) ? k.abw : k.abh;
    float *sc = k.SZc + ((size_t)f * NX + ix) * 3;
    float *sp = k.SZp + ((size_t)f * NX + ix) * 3;
    const float c0 = sc[0], c1 = sc[1], c2 = sc[2], p0 = sp[0], p1 = sp[1], p2 = sp[2];
    const float f1 = lds[(f * 2 + 1) * E + q], f2 = lds[(f * 2 + 0) * E + q];
    sp[0] = c0; sp[1] = c1; sp[2] = c2;
    if (upd) {
      const float n = cf[0] * f1 + cf[1] * f2 + cf[2] * c0 + cf[3] * c1 + cf[4] * c2 + cf[5] * p0 +
                      cf[6] * p1 + cf[7] * p2;
      FA(F[f], i, lst) = n;
      sc[0] = n; sc[1] = f1; sc[2] = f2;
    } else {
      sc[0] = FA(F[f], i, lst); sc[1] = f1; sc[2] = f2;
    }
  }
}

// Block b of a velocity launch over split sp: the edge blocks (fold), then
// nrest per-cell blocks, then nruns run blocks.
template <int NR = kRun>
__device__ __forceinline__ void velocity_block(const KParams &k, const Split &sp, int b, int nrest,
                                               int nruns, float *lds) {
  if (b < sp.etop) {
    top_block(k, sp, b, lds);
    return;
  }
  b -= sp.etop;
  if (b < sp.ebot) {
    bottom_block(k, b, lds);
    return;
  }
  b -= sp.ebot;
  if (b < nrest) {
    int i, l;
    if (rest_cell(sp, b * kBlock + threadIdx.x, &i, &l)) velocity_cell(k, i, l);
    return;
  }
  int tile;
  const int r = run_rect(sp, xcd_tile(b - nrest, nruns), &tile);
  velocity_run<NR>(k, sp.qx0[r], sp.qxl[r], sp.qntx[r], sp.qz0[r], sp.qz1[r], tile, threadIdx.x >> 6,
                     threadIdx.x & 63);
}

}  // namespace bqk

using namespace bqk;
