/*
 * biot_oracle.c - CPU restatement of the reference Biot time loop.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker the HIP path is compared
 * against; it is never linked into, loaded by or called from the product
 * library. Only tests/, the smoke check and the CPU-baseline leg of the bench
 * may load it.
 *
 * It restates, in the reference's own loop order and expression order, the
 * body of TIME_LOOP (time_loop.F90:63-206) for the serial build: the row
 * drivers HPLANE_* and the 1-D row kernels STRAIN_*, STRESS_*, INH_*, STO_*,
 * including the module-array temporaries (EXX..QZZ, QXP, QZP) the reference
 * keeps. Each function names the reference routine and lines it follows.
 *
 * Compiled twice: REAL=float (the reference's PP=SP build, which it must match
 * bit for bit) and REAL=double (the FP64 oracle). Both builds must be compiled
 * without floating-point contraction so every operation rounds as written.
 * The interior rows of each half-step are independent of one another and are
 * shared out over OpenMP threads (one_step); each cell's expression is the
 * same, so the results do not depend on the thread count.
 *
 * Grid convention (aux1.F90:36-46, domain_decomposition.F90:146-154): x index
 * i in [PMLL, PMLXH] = [1-T, MX+T], z index l in [0, PMLZH]; every full-grid
 * array is stored x-fastest, element (i,l) at l*NX + (i-PMLL).
 *
 * Parity status: pinned bit-exactly against snapshots of the reference built
 * from its own sources (oracle/Makefile -> oracle/_ref/fd2d), see
 * tests/golden/make_golden.py and tests/test_oracle_golden.py, including the
 * CPU-runnable workload at its full size (400 x 400 cells, 1500 steps) by the
 * digests of every file the reference writes. The surface
 * rows without stress imaging (HPLANE_0/1_* non-imaging branches, with the
 * AFDA_ZZ / AFDA_XZ options as run-time bits) follow the subroutine bodies;
 * the reference itself faults in that variant (its INH_QZ_2ND_0 interface
 * block does not match the subroutine), so parity against the reference is
 * unpinned there and each surface stencil is pinned by polynomial exactness
 * instead (tests/test_oracle_surface.py).
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef REAL
#define REAL float
#endif
typedef REAL real;

#define EXPORT __attribute__((visibility("default")))

/* Field ids shared with the product ABI. */
enum { F_U = 0, F_W, F_QX, F_QZ, F_TXX, F_TZZ, F_TXZ, F_P, F_QXP, F_QZP, F_NFIELDS };

/* .MO table order (read_model_par.F90:52-60). */
enum {
  T_MU = 0, T_SXX1, T_SXX2, T_SXX3, T_SZZ1, T_SZZ2, T_SZZ3, T_P1, T_P2, T_P3,
  T_REL1_U, T_REL1_W, T_REL2_U, T_REL2_W, T_REL3_U, T_REL3_W,
  T_REL1_QX, T_REL1_QZ, T_REL2_QX, T_REL2_QZ, T_REL3_QX, T_REL3_QZ,
  T_AUXW, T_AUXQZ, T_N
};

typedef struct oracle_desc {
  int32_t mx, mz, tpml, stress_imaging, lin_x, lin_z, ktbo, mt2, jmnum, variant;
  double h, dt;
} oracle_desc;

/* variant bits: the reference's compile-time options - the one-sided stencils
 * of the non-imaging free surface (strain_uw_zz_0.F90:52, strain_q_zz_0.F90:47,
 * strain_uw_xz_1.F90:83) and the second-order build SECOND (mod_auxil.F90:15-21:
 * A = 0, B = 1; inh_qx_int_0.F90:40-47, inh_v_uv_int_0.F90:40-48) */
enum { V_AFDA_ZZ = 1, V_AFDA_XZ = 2, V_SECOND = 4 };

typedef struct {
  /* sizes (aux1.F90:36-46) */
  int MX, MZ, T, PMLL, PMLXH, PMLZH, NX, NZ, MX1;
  int LIN_X, LIN_Z, SI, AFDA_ZZ, AFDA_XZ, SECOND, KTBO, MT2, JMNUM, next_step;
  real H, DT, A, B, AH, BH, CH;
  int32_t *JM;
  real *tab[T_N];
  real *f[F_NFIELDS];
  /* row temporaries of MODULE STRAINFIELD (mod_strainfield.F90:7-9) */
  real *EXX, *EZZ, *EXZ, *QXX, *QZZ;
  /* per-thread copies of the row temporaries for the interior rows, which
   * the OpenMP loops of one_step share out: 5 x NX reals per thread */
  real *rowpool;
  int nrowpool;
  /* PML coefficients (mod_pml.F90:7-19): XL/XH are (0:T[+1], NZ), ZH (0:T[+1], NX) */
  real *ODA_XL, *ODB_XL, *ODC_XL, *ODD_XL, *ODA_XH, *ODB_XH, *ODC_XH, *ODD_XH;
  real *ODA_ZH, *ODB_ZH, *ODC_ZH, *ODD_ZH;
  real *ODA2_XL, *ODB2_XL, *ODC2_XL, *ODD2_XL, *ODA2_XH, *ODB2_XH, *ODC2_XH, *ODD2_XH;
  real *ODA2_ZH, *ODB2_ZH, *ODC2_ZH, *ODD2_ZH;
  /* PML memories (mod_pml.F90:21-31, shapes from ppml.F90:88-112) */
  real *PZXL, *PZXH, *PXXL, *PXXH, *OXXL, *OXXH, *PXZH, *PZZH, *OZZH;
  real *RXXL, *RXXH, *SXXL, *SXXH, *GXXL, *GXXH, *HXXL, *HXXH;
  real *RZXL, *RZXH, *GZXL, *GZXH;
  real *RXZH, *GXZH, *RZZH, *SZZH, *GZZH, *HZZH;
  /* bottom non-reflecting history (mod_nr_bound.F90:33-37), [NX][3] */
  real *SZ_U, *SZ_W, *SZ_QX, *SZ_QZ, *SZ_UP, *SZ_WP, *SZ_QXP, *SZ_QZP;
  real ABW[8], ABH[8];
  /* source (auxil SOURTF_A(2,MT2)) and its previous values */
  real *SOURTF;
  real SOURSP, SOURFP;
  /* receivers */
  int MR;
  int *IRECX, *LRECX, *IRECZ, *LRECZ;
  real *traces; /* [MT2][MR][6] */
  char err[256];
} orc;

/* ----------------------------------------------------------------- helpers */
#define IXF(o, i, l) ((size_t)(l) * (size_t)(o)->NX + (size_t)((i) - (o)->PMLL))
#define FLD(o, k, i, l) ((o)->f[k][IXF(o, i, l)])
#define UM(i, l) FLD(o, F_U, i, l)
#define WM(i, l) FLD(o, F_W, i, l)
#define QXF(i, l) FLD(o, F_QX, i, l)
#define QZF(i, l) FLD(o, F_QZ, i, l)
#define QXPF(i, l) FLD(o, F_QXP, i, l)
#define QZPF(i, l) FLD(o, F_QZP, i, l)
#define TXXF(i, l) FLD(o, F_TXX, i, l)
#define TZZF(i, l) FLD(o, F_TZZ, i, l)
#define TXZF(i, l) FLD(o, F_TXZ, i, l)
#define PRF(i, l) FLD(o, F_P, i, l)
#define JMF(i, l) (o->JM[IXF(o, i, l)] - 1)
#define TAB(k, jm) (o->tab[k][jm])
/* row temporaries indexed by i */
#define RT(a, i) ((a)[(i) - o->PMLL])
/* PML coefficients: XL1 (0:T, z), XL2 (0:T+1, z), ZH1 (0:T, x), ZH2 (0:T+1, x) */
#define XL1(a, j, l) ((a)[(size_t)(l) * (o->T + 1) + (j)])
#define XL2(a, j, l) ((a)[(size_t)(l) * (o->T + 2) + (j)])
#define ZH1(a, j, i) ((a)[(size_t)((i) - o->PMLL) * (o->T + 1) + (j)])
#define ZH2(a, j, i) ((a)[(size_t)((i) - o->PMLL) * (o->T + 2) + (j)])
/* memories: left strips start at PMLL, right strips at MX or MX-1 */
#define ML(a, w, i, l) ((a)[(size_t)(l) * (w) + ((i) - o->PMLL)])
#define MR0(a, w, i, l) ((a)[(size_t)(l) * (w) + ((i) - o->MX)])
#define MR1(a, w, i, l) ((a)[(size_t)(l) * (w) + ((i) - (o->MX - 1))])
#define MZB(a, i, ll) ((a)[(size_t)(ll) * o->NX + ((i) - o->PMLL)])
/* widths */
#define W_L1 (o->T + 1)   /* PMLL:1 */
#define W_L2 (o->T + 2)   /* PMLL:2 */
#define W_L3 (o->T + 3)   /* PMLL:3 */
#define W_R0 (o->T + 1)   /* MX:PMLXH */
#define W_R1 (o->T + 2)   /* MX-1:PMLXH */

static inline int imax(int a, int b) { return a > b ? a : b; }
static inline int imin(int a, int b) { return a < b ? a : b; }

static void *zalloc(size_t n) {
  void *p = calloc(n ? n : 1, 1);
  return p;
}

/* ======================================================= stress row kernels */

/* STRAIN_UW_XZ, strain_uw_xz.F90:61-103 */
static void strain_uw_xz(orc *o, int L, real *EXZ, real A, real B, real C) {
  const int T = o->T, MX = o->MX, MXT1 = MX - 1;
  for (int I = 2 - T; I <= 2; ++I) {
    real DZXA = C * (WM(I, L) - WM(I - 1, L));
    RT(EXZ, I) = C * (UM(I, L) - UM(I, L - 1)) +
                 XL2(o->ODA2_XL, 3 - I, L) *
                     (XL2(o->ODD2_XL, 3 - I, L) * DZXA + ML(o->PZXL, W_L2, I, L));
    ML(o->PZXL, W_L2, I, L) =
        XL2(o->ODB2_XL, 3 - I, L) * ML(o->PZXL, W_L2, I, L) + XL2(o->ODC2_XL, 3 - I, L) * DZXA;
  }
  for (int I = 3; I <= MXT1; ++I) {
    RT(EXZ, I) = A * (UM(I, L + 1) - UM(I, L - 2) + WM(I + 1, L) - WM(I - 2, L)) +
                 B * (UM(I, L) - UM(I, L - 1) + WM(I, L) - WM(I - 1, L));
  }
  for (int I = MX; I <= MX + T; ++I) {
    real DZXA = C * (WM(I, L) - WM(I - 1, L));
    RT(EXZ, I) = C * (UM(I, L) - UM(I, L - 1)) +
                 XL2(o->ODA2_XH, I - MXT1, L) *
                     (XL2(o->ODD2_XH, I - MXT1, L) * DZXA + MR0(o->PZXH, W_R0, I, L));
    MR0(o->PZXH, W_R0, I, L) = XL2(o->ODB2_XH, I - MXT1, L) * MR0(o->PZXH, W_R0, I, L) +
                               XL2(o->ODC2_XH, I - MXT1, L) * DZXA;
  }
}

/* STRAIN_UW_ZZ, strain_uw_zz.F90:49-60 (and STRAIN_Q_ZZ, strain_q_zz.F90:46-58) */
static void strain_zz(orc *o, int L, const real *F, real *E, real A, real B, real C) {
  const int T = o->T, MXT1 = o->MX - 1;
  real BC = B / C;
  for (int I = 2 - T; I <= MXT1 + T; ++I)
    RT(E, I) = (C * (F[IXF(o, I, L + 1)] - F[IXF(o, I, L)]));
  for (int I = 2; I <= MXT1; ++I)
    RT(E, I) = A * (F[IXF(o, I, L + 2)] - F[IXF(o, I, L - 1)]) + BC * RT(E, I);
}

/* STRAIN_UW_XX, strain_uw_xx.F90:60-107 (and STRAIN_Q_XX, strain_q_xx.F90:55-100) */
static void strain_xx(orc *o, int L, const real *F, real *E, real *ML_, real *MRh, real A,
                      real B, real C) {
  const int T = o->T, MX = o->MX, MXT1 = MX - 1;
#define FF(i) (F[IXF(o, i, L)])
  for (int I = 2 - T; I <= 1; ++I) {
    real DXXA = C * (FF(I + 1) - FF(I));
    RT(E, I) = XL1(o->ODA_XL, 2 - I, L) * (XL1(o->ODD_XL, 2 - I, L) * DXXA + ML(ML_, W_L1, I, L));
    ML(ML_, W_L1, I, L) =
        XL1(o->ODB_XL, 2 - I, L) * ML(ML_, W_L1, I, L) + XL1(o->ODC_XL, 2 - I, L) * DXXA;
  }
  RT(E, 2) = C * (FF(3) - FF(2));
  for (int I = 3; I <= MX - 2; ++I) RT(E, I) = A * (FF(I + 2) - FF(I - 1)) + B * (FF(I + 1) - FF(I));
  RT(E, MXT1) = C * (FF(MXT1 + 1) - FF(MXT1));
  for (int I = MX; I <= MXT1 + T; ++I) {
    real DXXA = C * (FF(I + 1) - FF(I));
    RT(E, I) = XL1(o->ODA_XH, I - MXT1, L) *
               (XL1(o->ODD_XH, I - MXT1, L) * DXXA + MR0(MRh, W_R0, I, L));
    MR0(MRh, W_R0, I, L) =
        XL1(o->ODB_XH, I - MXT1, L) * MR0(MRh, W_R0, I, L) + XL1(o->ODC_XH, I - MXT1, L) * DXXA;
  }
#undef FF
}

/* STRAIN_UW_ZZ_2ND_PML, strain_uw_zz_2nd_pml.F90:55-65 (and STRAIN_Q_ZZ_2ND_PML,
 * strain_q_zz_2nd_pml.F90:51-64) */
static void strain_zz_2nd_pml(orc *o, int L, int LL, const real *F, real *E, real *MZ,
                              real C) {
  const int T = o->T, MXT1 = o->MX - 1;
  for (int I = 2 - T; I <= MXT1 + T; ++I) {
    real DZZA = C * (F[IXF(o, I, L + 1)] - F[IXF(o, I, L)]);
    RT(E, I) = ZH1(o->ODA_ZH, LL, I) * (ZH1(o->ODD_ZH, LL, I) * DZZA + MZB(MZ, I, LL));
    MZB(MZ, I, LL) = ZH1(o->ODB_ZH, LL, I) * MZB(MZ, I, LL) + ZH1(o->ODC_ZH, LL, I) * DZZA;
  }
}

/* STRAIN_UW_XZ_2ND_PML, strain_uw_xz_2nd_pml.F90:69-121 */
static void strain_uw_xz_2nd_pml(orc *o, int L, int LL, real *EXZ, real C) {
  const int T = o->T, MX = o->MX, MXT1 = MX - 1;
  for (int I = 2 - T; I <= 2; ++I) {
    real DZXA = C * (WM(I, L) - WM(I - 1, L));
    real DXZA = C * (UM(I, L) - UM(I, L - 1));
    RT(EXZ, I) = ZH2(o->ODA2_ZH, LL, I) * (ZH2(o->ODD2_ZH, LL, I) * DXZA + MZB(o->PXZH, I, LL)) +
                 XL2(o->ODA2_XL, 3 - I, L) *
                     (XL2(o->ODD2_XL, 3 - I, L) * DZXA + ML(o->PZXL, W_L2, I, L));
    MZB(o->PXZH, I, LL) = ZH2(o->ODB2_ZH, LL, I) * MZB(o->PXZH, I, LL) + ZH2(o->ODC2_ZH, LL, I) * DXZA;
    ML(o->PZXL, W_L2, I, L) =
        XL2(o->ODB2_XL, 3 - I, L) * ML(o->PZXL, W_L2, I, L) + XL2(o->ODC2_XL, 3 - I, L) * DZXA;
  }
  for (int I = 3; I <= MXT1; ++I) {
    real DZXA = C * (WM(I, L) - WM(I - 1, L));
    real DXZA = C * (UM(I, L) - UM(I, L - 1));
    RT(EXZ, I) =
        ZH2(o->ODA2_ZH, LL, I) * (ZH2(o->ODD2_ZH, LL, I) * DXZA + MZB(o->PXZH, I, LL)) + DZXA;
    MZB(o->PXZH, I, LL) = ZH2(o->ODB2_ZH, LL, I) * MZB(o->PXZH, I, LL) + ZH2(o->ODC2_ZH, LL, I) * DXZA;
  }
  for (int I = MX; I <= MX + T; ++I) {
    real DZXA = C * (WM(I, L) - WM(I - 1, L));
    real DXZA = C * (UM(I, L) - UM(I, L - 1));
    RT(EXZ, I) = ZH2(o->ODA2_ZH, LL, I) * (ZH2(o->ODD2_ZH, LL, I) * DXZA + MZB(o->PXZH, I, LL)) +
                 XL2(o->ODA2_XH, I - MXT1, L) *
                     (XL2(o->ODD2_XH, I - MXT1, L) * DZXA + MR0(o->PZXH, W_R0, I, L));
    MZB(o->PXZH, I, LL) = ZH2(o->ODB2_ZH, LL, I) * MZB(o->PXZH, I, LL) + ZH2(o->ODC2_ZH, LL, I) * DXZA;
    MR0(o->PZXH, W_R0, I, L) = XL2(o->ODB2_XH, I - MXT1, L) * MR0(o->PZXH, W_R0, I, L) +
                               XL2(o->ODC2_XH, I - MXT1, L) * DZXA;
  }
}

/* STRESS_XZ, stress_xz.F90:40-47 (TAU_EPS is forced to 0, read_open.F90:171) */
static void stress_xz(orc *o, int L, const real *EXZ) {
  const int T = o->T;
  for (int I = 2 - T; I <= o->MX + T; ++I) {
    int jm = JMF(I, L);
    real TXZ = TAB(T_MU, jm) * RT(EXZ, I);
    TXZF(I, L) = TXZF(I, L) + (o->DT + (real)0) * TXZ;
  }
}

/* STRESS_NN, stress_nn.F90:49-61; with src != 0 STRESS_LIN_NN, stress_lin_nn.F90:49-70 */
static void stress_nn(orc *o, int L, const real *EXX, const real *EZZ, const real *QXX,
                      const real *QZZ, int src, real SOURS, real SOURF) {
  const int T = o->T;
  const real DT = o->DT + (real)0;
  for (int I = 2 - T; I <= o->MX - 1 + T; ++I) {
    int jm = JMF(I, L);
    real ex = RT(EXX, I), ez = RT(EZZ, I), qx = RT(QXX, I), qz = RT(QZZ, I);
    real TXX = TAB(T_SXX1, jm) * ex + TAB(T_SXX2, jm) * ez + TAB(T_SXX3, jm) * qx +
               TAB(T_SXX3, jm) * qz;
    real TZZ = TAB(T_SZZ1, jm) * ex + TAB(T_SZZ2, jm) * ez + TAB(T_SZZ3, jm) * qx +
               TAB(T_SZZ3, jm) * qz;
    real PRES =
        TAB(T_P1, jm) * ex + TAB(T_P2, jm) * ez + TAB(T_P3, jm) * qx + TAB(T_P3, jm) * qz;
    if (src && I == o->LIN_X) {
      TXXF(I, L) = TXXF(I, L) + DT * (TXX - ((SOURS - o->SOURSP) / o->DT));
      TZZF(I, L) = TZZF(I, L) + DT * (TZZ - ((SOURS - o->SOURSP) / o->DT));
      PRF(I, L) = PRF(I, L) + DT * (PRES + ((SOURF - o->SOURFP) / o->DT));
      o->SOURSP = SOURS;
      o->SOURFP = SOURF;
    } else {
      TXXF(I, L) = TXXF(I, L) + DT * TXX;
      TZZF(I, L) = TZZF(I, L) + DT * TZZ;
      PRF(I, L) = PRF(I, L) + DT * PRES;
    }
  }
}

typedef struct { real *EXX, *EZZ, *EXZ, *QXX, *QZZ; } rowtmp;

/* HPLANE_IC_STRESS, hplane_ic_stress.F90:48-117; HPLANE_LIN_STRESS,
 * hplane_lin_stress.F90:52-127 when src != 0 */
static void hplane_ic_stress(orc *o, int L, rowtmp *t, int src, real S, real F) {
  strain_uw_xz(o, L, t->EXZ, o->AH, o->BH, o->CH);
  strain_zz(o, L, o->f[F_W], t->EZZ, o->AH, o->BH, o->CH);
  strain_zz(o, L, o->f[F_QZ], t->QZZ, o->AH, o->BH, o->CH);
  strain_xx(o, L, o->f[F_U], t->EXX, o->PXXL, o->PXXH, o->AH, o->BH, o->CH);
  strain_xx(o, L, o->f[F_QX], t->QXX, o->OXXL, o->OXXH, o->AH, o->BH, o->CH);
  stress_xz(o, L, t->EXZ);
  stress_nn(o, L, t->EXX, t->EZZ, t->QXX, t->QZZ, src, S, F);
}

/* HPLANE_MZ1_STRESS, hplane_mz1_stress.F90:50-130 */
static void hplane_mz1_stress(orc *o, rowtmp *t) {
  int L = o->MZ - 1, LL = 0;
  strain_uw_xz(o, L, t->EXZ, o->AH, o->BH, o->CH);
  strain_zz_2nd_pml(o, L, LL, o->f[F_W], t->EZZ, o->PZZH, o->CH);
  strain_zz_2nd_pml(o, L, LL, o->f[F_QZ], t->QZZ, o->OZZH, o->CH);
  strain_xx(o, L, o->f[F_U], t->EXX, o->PXXL, o->PXXH, o->AH, o->BH, o->CH);
  strain_xx(o, L, o->f[F_QX], t->QXX, o->OXXL, o->OXXH, o->AH, o->BH, o->CH);
  stress_xz(o, L, t->EXZ);
  stress_nn(o, L, t->EXX, t->EZZ, t->QXX, t->QZZ, 0, 0, 0);
}

/* HPLANE_PML_STRESS, hplane_pml_stress.F90:53-134 */
static void hplane_pml_stress(orc *o, int L, int LL, rowtmp *t) {
  strain_uw_xz_2nd_pml(o, L, LL, t->EXZ, o->CH);
  strain_zz_2nd_pml(o, L, LL, o->f[F_W], t->EZZ, o->PZZH, o->CH);
  strain_zz_2nd_pml(o, L, LL, o->f[F_QZ], t->QZZ, o->OZZH, o->CH);
  strain_xx(o, L, o->f[F_U], t->EXX, o->PXXL, o->PXXH, o->AH, o->BH, o->CH);
  strain_xx(o, L, o->f[F_QX], t->QXX, o->OXXL, o->OXXH, o->AH, o->BH, o->CH);
  stress_xz(o, L, t->EXZ);
  stress_nn(o, L, t->EXX, t->EZZ, t->QXX, t->QZZ, 0, 0, 0);
}

/* HPLANE_PMLZH0_STRESS, hplane_pmlzh0_stress.F90:43-68 */
static void hplane_pmlzh0_stress(orc *o, rowtmp *t) {
  int L = o->MZ + o->T, LL = o->T + 1;
  strain_uw_xz_2nd_pml(o, L, LL, t->EXZ, o->CH);
  stress_xz(o, L, t->EXZ);
}

/* HPLANE_0_STRESS + HPLANE_1_STRESS, stress-imaging branch
 * (hplane_0_stress.F90:47-51, hplane_1_stress.F90:44-49) */
static void hplane_01_stress_si(orc *o) {
  for (int I = o->PMLL; I <= o->PMLXH; ++I) {
    TXZF(I, 0) = -TXZF(I, 4);
    TZZF(I, 0) = -TZZF(I, 3);
    PRF(I, 0) = -PRF(I, 3);
  }
  for (int I = o->PMLL; I <= o->PMLXH; ++I) {
    TXZF(I, 1) = -TXZF(I, 3);
    TZZF(I, 1) = -TZZF(I, 2);
    PRF(I, 1) = -PRF(I, 2);
    TXZF(I, 2) = 0;
  }
}

/* STRAIN_UW_ZZ_0, strain_uw_zz_0.F90:45-62 (and STRAIN_Q_ZZ_0, strain_q_zz_0.F90:40-57):
 * row 0 without stress imaging, one-sided 5-point form with AFDA_ZZ */
static void strain_zz_0(orc *o, const real *F, real *E, real C) {
  const int T = o->T, MXT1 = o->MX - 1;
  for (int I = 2 - T; I <= MXT1 + T; ++I) RT(E, I) = (C * (F[IXF(o, I, 1)] - F[IXF(o, I, 0)]));
  if (o->AFDA_ZZ)
    for (int I = 2; I <= MXT1; ++I)
      RT(E, I) = C * (-((real)11 / (real)12 * F[IXF(o, I, 0)]) + (real)17 / (real)24 * F[IXF(o, I, 1)] +
                      (real)3 / (real)8 * F[IXF(o, I, 2)] - (real)5 / (real)24 * F[IXF(o, I, 3)] +
                      (real)1 / (real)24 * F[IXF(o, I, 4)]);
}

/* STRAIN_UW_XZ_1, strain_uw_xz_1.F90:62-114: row 1 without stress imaging; the
 * z-derivative of U is the 2-point one (U(i,0) is the surface row), or a
 * one-sided 4-point one with AFDA_XZ */
static void strain_uw_xz_1(orc *o, real *EXZ, real A, real B, real C) {
  const int T = o->T, MX = o->MX, MXT1 = MX - 1, L = 1, L1 = 0;
  for (int I = 2 - T; I <= 2; ++I) {
    real DZXA = C * (WM(I, L) - WM(I - 1, L));
    RT(EXZ, I) = C * (UM(I, L) - UM(I, L1)) +
                 XL2(o->ODA2_XL, 3 - I, L) *
                     (XL2(o->ODD2_XL, 3 - I, L) * DZXA + ML(o->PZXL, W_L2, I, L));
    ML(o->PZXL, W_L2, I, L) =
        XL2(o->ODB2_XL, 3 - I, L) * ML(o->PZXL, W_L2, I, L) + XL2(o->ODC2_XL, 3 - I, L) * DZXA;
  }
  for (int I = 3; I <= MX - 1; ++I) {
    if (o->AFDA_XZ) {
      real DUIKM1 = -(A * (WM(I + 1, 0) - WM(I - 2, 0))) - B * (WM(I, 0) - WM(I - 1, 0));
      RT(EXZ, I) = A * (WM(I + 1, L) - WM(I - 2, L)) + B * (WM(I, L) - WM(I - 1, L)) -
                   (real)1 / (real)22 * DUIKM1 +
                   C * (-((real)577 / (real)528 * UM(I, 0)) + (real)201 / (real)176 * UM(I, 1) -
                        (real)9 / (real)176 * UM(I, 2) + (real)1 / (real)528 * UM(I, 3));
    } else {
      RT(EXZ, I) = A * (WM(I + 1, L) - WM(I - 2, L)) + B * (WM(I, L) - WM(I - 1, L)) +
                   C * (UM(I, 1) - UM(I, 0));
    }
  }
  for (int I = MX; I <= MX + T; ++I) {
    real DZXA = C * (WM(I, L) - WM(I - 1, L));
    RT(EXZ, I) = C * (UM(I, L) - UM(I, L1)) +
                 XL2(o->ODA2_XH, I - MXT1, L) *
                     (XL2(o->ODD2_XH, I - MXT1, L) * DZXA + MR0(o->PZXH, W_R0, I, L));
    MR0(o->PZXH, W_R0, I, L) = XL2(o->ODB2_XH, I - MXT1, L) * MR0(o->PZXH, W_R0, I, L) +
                               XL2(o->ODC2_XH, I - MXT1, L) * DZXA;
  }
}

/* HPLANE_0_STRESS without stress imaging, hplane_0_stress.F90:57-119: EXZ = 0 */
static void hplane_0_stress(orc *o, rowtmp *t) {
  for (int I = o->PMLL; I <= o->PMLXH; ++I) RT(t->EXZ, I) = 0;
  strain_zz_0(o, o->f[F_W], t->EZZ, o->CH);
  strain_zz_0(o, o->f[F_QZ], t->QZZ, o->CH);
  strain_xx(o, 0, o->f[F_U], t->EXX, o->PXXL, o->PXXH, o->AH, o->BH, o->CH);
  strain_xx(o, 0, o->f[F_QX], t->QXX, o->OXXL, o->OXXH, o->AH, o->BH, o->CH);
  stress_xz(o, 0, t->EXZ);
  stress_nn(o, 0, t->EXX, t->EZZ, t->QXX, t->QZZ, 0, 0, 0);
}

/* HPLANE_1_STRESS without stress imaging, hplane_1_stress.F90:51-126 */
static void hplane_1_stress(orc *o, rowtmp *t) {
  strain_uw_xz_1(o, t->EXZ, o->AH, o->BH, o->CH);
  strain_zz(o, 1, o->f[F_W], t->EZZ, o->AH, o->BH, o->CH);
  strain_zz(o, 1, o->f[F_QZ], t->QZZ, o->AH, o->BH, o->CH);
  strain_xx(o, 1, o->f[F_U], t->EXX, o->PXXL, o->PXXH, o->AH, o->BH, o->CH);
  strain_xx(o, 1, o->f[F_QX], t->QXX, o->OXXL, o->OXXH, o->AH, o->BH, o->CH);
  stress_xz(o, 1, t->EXZ);
  stress_nn(o, 1, t->EXX, t->EZZ, t->QXX, t->QZZ, 0, 0, 0);
}

/* ===================================================== velocity row kernels */

/* precomputed per-material friction factors are NOT used here: the oracle
 * evaluates EXP(-REL3) and (REL2/REL3)*(1-EXP(-REL3)) per cell as the
 * reference does (inh_qx_int.F90:44, inh_v_uv_int.F90:46). */
static inline real rexp(real x) {
#if defined(ORACLE_DOUBLE)
  return exp(x);
#else
  return expf(x);
#endif
}

/* INH_QX_INT, inh_qx_int.F90:40-54 */
static void inh_qx_int(orc *o, int L) {
  const real A = o->A, B = o->B;
  for (int I = 4; I <= o->MX - 2; ++I) {
    int jm = JMF(I, L);
    QXF(I, L) = rexp(-TAB(T_REL3_QX, jm)) * QXF(I, L) -
                TAB(T_REL1_QX, jm) *
                    (A * (TXXF(I + 1, L) - TXXF(I - 2, L) + TXZF(I, L + 2) - TXZF(I, L - 1)) +
                     B * (TXXF(I, L) - TXXF(I - 1, L) + TXZF(I, L + 1) - TXZF(I, L))) -
                TAB(T_REL2_QX, jm) * (A * (PRF(I + 1, L) - PRF(I - 2, L)) + B * (PRF(I, L) - PRF(I - 1, L)));
  }
}

/* INH_QZ_INT, inh_qz_int.F90:41-55 */
static void inh_qz_int(orc *o, int L) {
  const real A = o->A, B = o->B;
  for (int I = 3; I <= o->MX - 2; ++I) {
    int jm = JMF(I, L);
    QZF(I, L) = rexp(-TAB(T_REL3_QZ, jm)) * QZF(I, L) -
                TAB(T_REL1_QZ, jm) *
                    (A * (TZZF(I, L + 1) - TZZF(I, L - 2) + TXZF(I + 2, L) - TXZF(I - 1, L)) +
                     B * (TZZF(I, L) - TZZF(I, L - 1) + TXZF(I + 1, L) - TXZF(I, L))) -
                TAB(T_REL2_QZ, jm) * (A * (PRF(I, L + 1) - PRF(I, L - 2)) + B * (PRF(I, L) - PRF(I, L - 1)));
  }
}

/* INH_V_UV_INT, inh_v_uv_int.F90:41-56 */
static void inh_v_uv_int(orc *o, int L) {
  const real A = o->A, B = o->B;
  for (int I = 4; I <= o->MX - 2; ++I) {
    int jm = JMF(I, L);
    UM(I, L) = UM(I, L) +
               (TAB(T_REL2_U, jm) / TAB(T_REL3_QX, jm)) * ((real)1 - rexp(-TAB(T_REL3_QX, jm))) *
                   QXPF(I, L) +
               TAB(T_REL1_U, jm) *
                   (A * (TXXF(I + 1, L) - TXXF(I - 2, L) + TXZF(I, L + 2) - TXZF(I, L - 1)) +
                    B * (TXXF(I, L) - TXXF(I - 1, L) + TXZF(I, L + 1) - TXZF(I, L))) +
               TAB(T_REL3_U, jm) * (A * (PRF(I + 1, L) - PRF(I - 2, L)) + B * (PRF(I, L) - PRF(I - 1, L)));
  }
}

/* INH_V_W_INT, inh_v_w_int.F90:43-58 */
static void inh_v_w_int(orc *o, int L) {
  const real A = o->A, B = o->B;
  for (int I = 3; I <= o->MX - 2; ++I) {
    int jm = JMF(I, L);
    WM(I, L) = WM(I, L) +
               (TAB(T_REL2_W, jm) / TAB(T_REL3_QZ, jm)) * ((real)1 - rexp(-TAB(T_REL3_QZ, jm))) *
                   QZPF(I, L) +
               TAB(T_REL1_W, jm) *
                   (A * (TZZF(I, L + 1) - TZZF(I, L - 2) + TXZF(I + 2, L) - TXZF(I - 1, L)) +
                    B * (TZZF(I, L) - TZZF(I, L - 1) + TXZF(I + 1, L) - TXZF(I, L))) +
               TAB(T_REL3_W, jm) * (A * (PRF(I, L + 1) - PRF(I, L - 2)) + B * (PRF(I, L) - PRF(I, L - 1)));
  }
}

/* The left and right x-PML column ranges of the velocity kernels. */
#define LEFT_UQX_LO (3 - o->T)
#define LEFT_UQX_HI 3
#define LEFT_WQZ_LO (2 - o->T)
#define LEFT_WQZ_HI 2
#define RIGHT_LO (o->MX1)
#define RIGHT_HI (o->MX1 + o->T)

/* INH_QX_2ND, inh_qx_2nd.F90:54-103 */
static void inh_qx_2nd(orc *o, int L) {
  for (int side = 0; side < 2; ++side) {
    int lo = side ? RIGHT_LO : LEFT_UQX_LO, hi = side ? RIGHT_HI : LEFT_UQX_HI;
    for (int I = lo; I <= hi; ++I) {
      int jm = JMF(I, L);
      int j = side ? I - o->MX1 : 3 - I;
      real oda = side ? XL2(o->ODA2_XH, j, L) : XL2(o->ODA2_XL, j, L);
      real odb = side ? XL2(o->ODB2_XH, j, L) : XL2(o->ODB2_XL, j, L);
      real odc = side ? XL2(o->ODC2_XH, j, L) : XL2(o->ODC2_XL, j, L);
      real odd = side ? XL2(o->ODD2_XH, j, L) : XL2(o->ODD2_XL, j, L);
      real *G = side ? &MR1(o->GXXH, W_R1, I, L) : &ML(o->GXXL, W_L3, I, L);
      real *Hh = side ? &MR1(o->HXXH, W_R1, I, L) : &ML(o->HXXL, W_L3, I, L);
      real DXXXA = TXXF(I, L) - TXXF(I - 1, L);
      real PXXXA = PRF(I, L) - PRF(I - 1, L);
      QXF(I, L) = rexp(-TAB(T_REL3_QX, jm)) * QXF(I, L) -
                  TAB(T_REL1_QX, jm) * (oda * (odd * DXXXA + *G) + TXZF(I, L + 1) - TXZF(I, L)) -
                  TAB(T_REL2_QX, jm) * (oda * (odd * PXXXA + *Hh));
      *G = odb * *G + odc * DXXXA;
      *Hh = odb * *Hh + odc * PXXXA;
    }
  }
}

/* INH_QZ_2ND, inh_qz_2nd.F90:49-88 */
static void inh_qz_2nd(orc *o, int L) {
  for (int side = 0; side < 2; ++side) {
    int lo = side ? RIGHT_LO : LEFT_WQZ_LO, hi = side ? RIGHT_HI : LEFT_WQZ_HI;
    for (int I = lo; I <= hi; ++I) {
      int jm = JMF(I, L);
      int j = side ? I - o->MX1 : 2 - I;
      real oda = side ? XL1(o->ODA_XH, j, L) : XL1(o->ODA_XL, j, L);
      real odb = side ? XL1(o->ODB_XH, j, L) : XL1(o->ODB_XL, j, L);
      real odc = side ? XL1(o->ODC_XH, j, L) : XL1(o->ODC_XL, j, L);
      real odd = side ? XL1(o->ODD_XH, j, L) : XL1(o->ODD_XL, j, L);
      real *G = side ? &MR1(o->GZXH, W_R1, I, L) : &ML(o->GZXL, W_L2, I, L);
      real DXZXA = TXZF(I + 1, L) - TXZF(I, L);
      QZF(I, L) = rexp(-TAB(T_REL3_QZ, jm)) * QZF(I, L) -
                  TAB(T_REL1_QZ, jm) * (TZZF(I, L) - TZZF(I, L - 1) + oda * (odd * DXZXA + *G)) -
                  TAB(T_REL2_QZ, jm) * (PRF(I, L) - PRF(I, L - 1));
      *G = odb * *G + odc * DXZXA;
    }
  }
}

/* INH_V_U_2ND, inh_v_u_2nd.F90:56-104 */
static void inh_v_u_2nd(orc *o, int L) {
  for (int side = 0; side < 2; ++side) {
    int lo = side ? RIGHT_LO : LEFT_UQX_LO, hi = side ? RIGHT_HI : LEFT_UQX_HI;
    for (int I = lo; I <= hi; ++I) {
      int jm = JMF(I, L);
      int j = side ? I - o->MX1 : 3 - I;
      real oda = side ? XL2(o->ODA2_XH, j, L) : XL2(o->ODA2_XL, j, L);
      real odb = side ? XL2(o->ODB2_XH, j, L) : XL2(o->ODB2_XL, j, L);
      real odc = side ? XL2(o->ODC2_XH, j, L) : XL2(o->ODC2_XL, j, L);
      real odd = side ? XL2(o->ODD2_XH, j, L) : XL2(o->ODD2_XL, j, L);
      real *R = side ? &MR1(o->RXXH, W_R1, I, L) : &ML(o->RXXL, W_L3, I, L);
      real *S = side ? &MR1(o->SXXH, W_R1, I, L) : &ML(o->SXXL, W_L3, I, L);
      real DXXXA = TXXF(I, L) - TXXF(I - 1, L);
      real PXXXA = PRF(I, L) - PRF(I - 1, L);
      UM(I, L) = UM(I, L) +
                 (TAB(T_REL2_U, jm) / TAB(T_REL3_QX, jm)) * ((real)1 - rexp(-TAB(T_REL3_QX, jm))) *
                     QXPF(I, L) +
                 TAB(T_REL1_U, jm) * (oda * (odd * DXXXA + *R) + TXZF(I, L + 1) - TXZF(I, L)) +
                 TAB(T_REL3_U, jm) * (oda * (odd * PXXXA + *S));
      *R = odb * *R + odc * DXXXA;
      *S = odb * *S + odc * PXXXA;
    }
  }
}

/* INH_V_W_2ND, inh_v_w_2nd.F90:52-91 */
static void inh_v_w_2nd(orc *o, int L) {
  for (int side = 0; side < 2; ++side) {
    int lo = side ? RIGHT_LO : LEFT_WQZ_LO, hi = side ? RIGHT_HI : LEFT_WQZ_HI;
    for (int I = lo; I <= hi; ++I) {
      int jm = JMF(I, L);
      int j = side ? I - o->MX1 : 2 - I;
      real oda = side ? XL1(o->ODA_XH, j, L) : XL1(o->ODA_XL, j, L);
      real odb = side ? XL1(o->ODB_XH, j, L) : XL1(o->ODB_XL, j, L);
      real odc = side ? XL1(o->ODC_XH, j, L) : XL1(o->ODC_XL, j, L);
      real odd = side ? XL1(o->ODD_XH, j, L) : XL1(o->ODD_XL, j, L);
      real *R = side ? &MR1(o->RZXH, W_R1, I, L) : &ML(o->RZXL, W_L2, I, L);
      real DXZXA = TXZF(I + 1, L) - TXZF(I, L);
      WM(I, L) = WM(I, L) +
                 (TAB(T_REL2_W, jm) / TAB(T_REL3_QZ, jm)) * ((real)1 - rexp(-TAB(T_REL3_QZ, jm))) *
                     QZPF(I, L) +
                 TAB(T_REL1_W, jm) * (TZZF(I, L) - TZZF(I, L - 1) + oda * (odd * DXZXA + *R)) +
                 TAB(T_REL3_W, jm) * (PRF(I, L) - PRF(I, L - 1));
      *R = odb * *R + odc * DXZXA;
    }
  }
}

/* STO_QXP, sto_qxp.F90:24-28; STO_QZP, sto_qzp.F90:24-28 */
static void sto_qp(orc *o, int L) {
  for (int I = 3 - o->T; I <= o->MX1 + o->T; ++I) QXPF(I, L) = QXF(I, L);
  for (int I = 2 - o->T; I <= o->MX1 + o->T; ++I) QZPF(I, L) = QZF(I, L);
}

/* HPLANE_IC_VELOCITY, hplane_ic_velocity.F90:59-202 (STO_BOR stores have no reader) */
static void hplane_ic_velocity(orc *o, int L) {
  inh_qx_int(o, L);
  inh_qz_int(o, L);
  inh_v_uv_int(o, L);
  inh_v_w_int(o, L);
  inh_qx_2nd(o, L);
  inh_qz_2nd(o, L);
  inh_v_u_2nd(o, L);
  inh_v_w_2nd(o, L);
  sto_qp(o, L);
}

/* INH_QX_INT_2ND_PML, inh_qx_int_2nd_pml.F90:60-143 */
static void inh_qx_int_2nd_pml(orc *o, int L, int LL) {
  const int MX1 = o->MX1;
  for (int I = 3 - o->T; I <= MX1 + o->T; ++I) {
    int region = I <= 3 ? 0 : (I <= MX1 - 1 ? 1 : 2);
    int jm = JMF(I, L);
    real DXXXA = TXXF(I, L) - TXXF(I - 1, L);
    real DXZZA = TXZF(I, L + 1) - TXZF(I, L);
    real e = rexp(-TAB(T_REL3_QX, jm));
    real *GZ = &MZB(o->GXZH, I, LL);
    real zo_a = ZH1(o->ODA_ZH, LL, I), zo_b = ZH1(o->ODB_ZH, LL, I);
    real zo_c = ZH1(o->ODC_ZH, LL, I), zo_d = ZH1(o->ODD_ZH, LL, I);
    if (region == 1) {
      QXF(I, L) = e * QXF(I, L) - TAB(T_REL1_QX, jm) * (zo_a * (zo_d * DXZZA + *GZ) + DXXXA) -
                  TAB(T_REL2_QX, jm) * (PRF(I, L) - PRF(I - 1, L));
      *GZ = zo_b * *GZ + zo_c * DXZZA;
    } else {
      int j = region ? I - MX1 : 3 - I;
      real oda = region ? XL2(o->ODA2_XH, j, L) : XL2(o->ODA2_XL, j, L);
      real odb = region ? XL2(o->ODB2_XH, j, L) : XL2(o->ODB2_XL, j, L);
      real odc = region ? XL2(o->ODC2_XH, j, L) : XL2(o->ODC2_XL, j, L);
      real odd = region ? XL2(o->ODD2_XH, j, L) : XL2(o->ODD2_XL, j, L);
      real *G = region ? &MR1(o->GXXH, W_R1, I, L) : &ML(o->GXXL, W_L3, I, L);
      real *Hh = region ? &MR1(o->HXXH, W_R1, I, L) : &ML(o->HXXL, W_L3, I, L);
      real PXXXA = PRF(I, L) - PRF(I - 1, L);
      QXF(I, L) = e * QXF(I, L) - TAB(T_REL1_QX, jm) * (oda * (odd * DXXXA + *G)) -
                  TAB(T_REL1_QX, jm) * (zo_a * (zo_d * DXZZA + *GZ)) -
                  TAB(T_REL2_QX, jm) * (oda * (odd * PXXXA + *Hh));
      *G = odb * *G + odc * DXXXA;
      *Hh = odb * *Hh + odc * PXXXA;
      *GZ = zo_b * *GZ + zo_c * DXZZA;
    }
  }
}

/* INH_V_U_INT_2ND_PML, inh_v_u_int_2nd_pml.F90:62-143 */
static void inh_v_u_int_2nd_pml(orc *o, int L, int LL) {
  const int MX1 = o->MX1;
  for (int I = 3 - o->T; I <= MX1 + o->T; ++I) {
    int region = I <= 3 ? 0 : (I <= MX1 - 1 ? 1 : 2);
    int jm = JMF(I, L);
    real DXXXA = TXXF(I, L) - TXXF(I - 1, L);
    real DXZZA = TXZF(I, L + 1) - TXZF(I, L);
    real PXXXA = PRF(I, L) - PRF(I - 1, L);
    real fr = (TAB(T_REL2_U, jm) / TAB(T_REL3_QX, jm)) * ((real)1 - rexp(-TAB(T_REL3_QX, jm)));
    real *RZ = &MZB(o->RXZH, I, LL);
    real zo_a = ZH1(o->ODA_ZH, LL, I), zo_b = ZH1(o->ODB_ZH, LL, I);
    real zo_c = ZH1(o->ODC_ZH, LL, I), zo_d = ZH1(o->ODD_ZH, LL, I);
    if (region == 1) {
      UM(I, L) = UM(I, L) + fr * QXPF(I, L) +
                 TAB(T_REL1_U, jm) * (DXXXA + zo_a * (zo_d * DXZZA + *RZ)) +
                 TAB(T_REL3_U, jm) * PXXXA;
      *RZ = zo_b * *RZ + zo_c * DXZZA;
    } else {
      int j = region ? I - MX1 : 3 - I;
      real oda = region ? XL2(o->ODA2_XH, j, L) : XL2(o->ODA2_XL, j, L);
      real odb = region ? XL2(o->ODB2_XH, j, L) : XL2(o->ODB2_XL, j, L);
      real odc = region ? XL2(o->ODC2_XH, j, L) : XL2(o->ODC2_XL, j, L);
      real odd = region ? XL2(o->ODD2_XH, j, L) : XL2(o->ODD2_XL, j, L);
      real *R = region ? &MR1(o->RXXH, W_R1, I, L) : &ML(o->RXXL, W_L3, I, L);
      real *S = region ? &MR1(o->SXXH, W_R1, I, L) : &ML(o->SXXL, W_L3, I, L);
      UM(I, L) = UM(I, L) + fr * QXPF(I, L) +
                 TAB(T_REL1_U, jm) * (oda * (odd * DXXXA + *R) + zo_a * (zo_d * DXZZA + *RZ)) +
                 TAB(T_REL3_U, jm) * (oda * (odd * PXXXA + *S));
      *R = odb * *R + odc * DXXXA;
      *S = odb * *S + odc * PXXXA;
      *RZ = zo_b * *RZ + zo_c * DXZZA;
    }
  }
}

/* INH_QZ_INT_2ND_PML, inh_qz_int_2nd_pml.F90:57-144 */
static void inh_qz_int_2nd_pml(orc *o, int L, int LL) {
  const int MX1 = o->MX1;
  for (int I = 2 - o->T; I <= MX1 + o->T; ++I) {
    int region = I <= 2 ? 0 : (I <= MX1 - 1 ? 1 : 2);
    int jm = JMF(I, L);
    real DXZXA = TXZF(I + 1, L) - TXZF(I, L);
    real DZZZA = TZZF(I, L) - TZZF(I, L - 1);
    real PZZZA = PRF(I, L) - PRF(I, L - 1);
    real e = rexp(-TAB(T_REL3_QZ, jm));
    real *GZ = &MZB(o->GZZH, I, LL), *HZ = &MZB(o->HZZH, I, LL);
    real za = ZH2(o->ODA2_ZH, LL, I), zb = ZH2(o->ODB2_ZH, LL, I);
    real zc = ZH2(o->ODC2_ZH, LL, I), zd = ZH2(o->ODD2_ZH, LL, I);
    if (region == 1) {
      QZF(I, L) = e * QZF(I, L) - TAB(T_REL1_QZ, jm) * (za * (zd * DZZZA + *GZ) + DXZXA) -
                  TAB(T_REL2_QZ, jm) * (za * (zd * PZZZA + *HZ));
      *GZ = zb * *GZ + zc * DZZZA;
      *HZ = zb * *HZ + zc * PZZZA;
    } else {
      int j = region ? I - MX1 : 2 - I;
      real oda = region ? XL1(o->ODA_XH, j, L) : XL1(o->ODA_XL, j, L);
      real odb = region ? XL1(o->ODB_XH, j, L) : XL1(o->ODB_XL, j, L);
      real odc = region ? XL1(o->ODC_XH, j, L) : XL1(o->ODC_XL, j, L);
      real odd = region ? XL1(o->ODD_XH, j, L) : XL1(o->ODD_XL, j, L);
      real *G = region ? &MR1(o->GZXH, W_R1, I, L) : &ML(o->GZXL, W_L2, I, L);
      QZF(I, L) = e * QZF(I, L) - TAB(T_REL1_QZ, jm) * (za * (zd * DZZZA + *GZ)) -
                  TAB(T_REL1_QZ, jm) * (oda * (odd * DXZXA + *G)) -
                  TAB(T_REL2_QZ, jm) * (za * (zd * PZZZA + *HZ));
      *G = odb * *G + odc * DXZXA;
      *GZ = zb * *GZ + zc * DZZZA;
      *HZ = zb * *HZ + zc * PZZZA;
    }
  }
}

/* INH_V_W_INT_2ND_PML, inh_v_w_int_2nd_pml.F90:61-146 */
static void inh_v_w_int_2nd_pml(orc *o, int L, int LL) {
  const int MX1 = o->MX1;
  for (int I = 2 - o->T; I <= MX1 + o->T; ++I) {
    int region = I <= 2 ? 0 : (I <= MX1 - 1 ? 1 : 2);
    int jm = JMF(I, L);
    real DXZXA = TXZF(I + 1, L) - TXZF(I, L);
    real DZZZA = TZZF(I, L) - TZZF(I, L - 1);
    real PZZZA = PRF(I, L) - PRF(I, L - 1);
    real fr = (TAB(T_REL2_W, jm) / TAB(T_REL3_QZ, jm)) * ((real)1 - rexp(-TAB(T_REL3_QZ, jm)));
    real *RZ = &MZB(o->RZZH, I, LL), *SZ = &MZB(o->SZZH, I, LL);
    real za = ZH2(o->ODA2_ZH, LL, I), zb = ZH2(o->ODB2_ZH, LL, I);
    real zc = ZH2(o->ODC2_ZH, LL, I), zd = ZH2(o->ODD2_ZH, LL, I);
    if (region == 1) {
      WM(I, L) = WM(I, L) + fr * QZPF(I, L) + TAB(T_REL1_W, jm) * (za * (zd * DZZZA + *RZ) + DXZXA) +
                 TAB(T_REL3_W, jm) * (za * (zd * PZZZA + *SZ));
      *RZ = zb * *RZ + zc * DZZZA;
      *SZ = zb * *SZ + zc * PZZZA;
    } else {
      int j = region ? I - MX1 : 2 - I;
      real oda = region ? XL1(o->ODA_XH, j, L) : XL1(o->ODA_XL, j, L);
      real odb = region ? XL1(o->ODB_XH, j, L) : XL1(o->ODB_XL, j, L);
      real odc = region ? XL1(o->ODC_XH, j, L) : XL1(o->ODC_XL, j, L);
      real odd = region ? XL1(o->ODD_XH, j, L) : XL1(o->ODD_XL, j, L);
      real *R = region ? &MR1(o->RZXH, W_R1, I, L) : &ML(o->RZXL, W_L2, I, L);
      WM(I, L) = WM(I, L) + fr * QZPF(I, L) +
                 TAB(T_REL1_W, jm) * (za * (zd * DZZZA + *RZ) + oda * (odd * DXZXA + *R)) +
                 TAB(T_REL3_W, jm) * (za * (zd * PZZZA + *SZ));
      *R = odb * *R + odc * DXZXA;
      *RZ = zb * *RZ + zc * DZZZA;
      *SZ = zb * *SZ + zc * PZZZA;
    }
  }
}

/* HPLANE_PML_VELOCITY, hplane_pml_velocity.F90:64-180 */
static void hplane_pml_velocity(orc *o, int L, int LL) {
  inh_qx_int_2nd_pml(o, L, LL);
  inh_v_u_int_2nd_pml(o, L, LL);
  inh_qz_int_2nd_pml(o, L, LL);
  inh_v_w_int_2nd_pml(o, L, LL);
  sto_qp(o, L);
}

/* HPLANE_1_VELOCITY stress-imaging branch: INH_V_QX_INT_1_SI
 * (inh_v_qx_int_1_si.F90:26-31), INH_W_QZ_INT_1_SI (inh_w_qz_int_1_si.F90:30-44),
 * then STO_QXP/STO_QZP of row 1 (hplane_1_velocity.F90:61-76,209-223) */
static void hplane_1_velocity_si(orc *o) {
  const int T = o->T, MXT = o->MX;
  for (int I = 3 - T; I <= MXT - 1 + T; ++I) {
    UM(I, 1) = UM(I, 2) + WM(I, 2) - WM(I - 1, 2);
    QXF(I, 1) = QXF(I, 2) + QZF(I, 2) - QZF(I - 1, 2);
  }
  for (int I = 3 - T; I <= MXT - 2 + T; ++I) {
    int jm = JMF(I, 0);
    WM(I, 1) = WM(I, 3) + TAB(T_AUXW, jm) * (UM(I + 1, 1) - UM(I, 1) + UM(I + 1, 2) - UM(I, 2));
    QZF(I, 1) = QZF(I, 3) + TAB(T_AUXQZ, jm) * (UM(I + 1, 1) - UM(I, 1) + UM(I + 1, 2) - UM(I, 2)) +
                (QXF(I + 1, 1) - QXF(I, 1) + QXF(I + 1, 2) - QXF(I, 2));
  }
  sto_qp(o, 1);
}

/* ---- rows 0 and 1 without stress imaging (hplane_0_velocity.F90:62-179,
 *      hplane_1_velocity.F90:79-207) */

/* lead + the z-derivative terms of U/QX on row 0 (TXZ(i,0) = 0 on the free
 * surface), added term by term after lead as the reference's expression does
 * (inh_qx_int_0.F90, inh_v_uv_int_0.F90) */
#define TXZ_Z0(lead, I)                                                                        \
  ((((lead) + (real)17 / (real)24 * TXZF(I, 1)) + (real)3 / (real)8 * TXZF(I, 2)) -             \
   (real)5 / (real)24 * TXZF(I, 3) + (real)1 / (real)24 * TXZF(I, 4))
/* one-sided z-derivatives on row 0 and row 1 (inh_qz_int_0.F90, inh_v_qz_int_1.F90) */
#define DZ0(F, I)                                                                              \
  ((real)35 / (real)8 * F(I, 0) - (real)35 / (real)24 * F(I, 1) + (real)21 / (real)40 * F(I, 2) - \
   (real)5 / (real)56 * F(I, 3))
#define DZ1(F, I)                                                                                \
  (-((real)31 / (real)24 * F(I, 0)) + (real)29 / (real)24 * F(I, 1) - (real)3 / (real)40 * F(I, 2) + \
   (real)1 / (real)168 * F(I, 3))
/* lead - 31/24 F0 + 29/24 F1 - 3/40 F2 + 1/168 F3, term by term (inh_v_qz_int_1.F90) */
#define DZ1L(lead, F, I)                                                                       \
  (((((lead) - (real)31 / (real)24 * F(I, 0)) + (real)29 / (real)24 * F(I, 1)) -               \
    (real)3 / (real)40 * F(I, 2)) + (real)1 / (real)168 * F(I, 3))

/* INH_QX_INT_0, inh_qx_int_0.F90:43-71; with SECOND the 2-point form of
 * :40-47, whose pressure term lacks a closing parenthesis in the reference
 * (so that build does not compile) and is read as (PRES(I,0) - PRES(I-1,0)) */
static void inh_qx_int_0(orc *o) {
  const real A = o->A, B = o->B;
  for (int I = 4; I <= o->MX - 2; ++I) {
    int jm = JMF(I, 0);
    if (o->SECOND) {
      QXF(I, 0) = rexp(-TAB(T_REL3_QX, jm)) * QXF(I, 0) -
                  TAB(T_REL1_QX, jm) * ((TXXF(I, 0) - TXXF(I - 1, 0)) + TXZF(I, 1)) -
                  TAB(T_REL2_QX, jm) * (PRF(I, 0) - PRF(I - 1, 0));
      continue;
    }
    QXF(I, 0) = rexp(-TAB(T_REL3_QX, jm)) * QXF(I, 0) -
                TAB(T_REL1_QX, jm) *
                    TXZ_Z0(A * (TXXF(I + 1, 0) - TXXF(I - 2, 0)) + B * (TXXF(I, 0) - TXXF(I - 1, 0)), I) -
                TAB(T_REL2_QX, jm) * (A * (PRF(I + 1, 0) - PRF(I - 2, 0)) + B * (PRF(I, 0) - PRF(I - 1, 0)));
  }
}

/* INH_V_UV_INT_0, inh_v_uv_int_0.F90:43-72; with SECOND the 2-point form of
 * :40-48 (same missing parenthesis) */
static void inh_v_uv_int_0(orc *o) {
  const real A = o->A, B = o->B;
  for (int I = 4; I <= o->MX - 2; ++I) {
    int jm = JMF(I, 0);
    if (o->SECOND) {
      UM(I, 0) = UM(I, 0) +
                 (TAB(T_REL2_U, jm) / TAB(T_REL3_QX, jm)) * ((real)1 - rexp(-TAB(T_REL3_QX, jm))) *
                     QXPF(I, 0) +
                 TAB(T_REL1_U, jm) * ((TXXF(I, 0) - TXXF(I - 1, 0)) + TXZF(I, 1)) +
                 TAB(T_REL3_U, jm) * (PRF(I, 0) - PRF(I - 1, 0));
      continue;
    }
    UM(I, 0) = UM(I, 0) +
               (TAB(T_REL2_U, jm) / TAB(T_REL3_QX, jm)) * ((real)1 - rexp(-TAB(T_REL3_QX, jm))) *
                   QXPF(I, 0) +
               TAB(T_REL1_U, jm) *
                   TXZ_Z0(A * (TXXF(I + 1, 0) - TXXF(I - 2, 0)) + B * (TXXF(I, 0) - TXXF(I - 1, 0)), I) +
               TAB(T_REL3_U, jm) * (A * (PRF(I + 1, 0) - PRF(I - 2, 0)) + B * (PRF(I, 0) - PRF(I - 1, 0)));
  }
}

/* INH_QZ_INT_0, inh_qz_int_0.F90:43-62 */
static void inh_qz_int_0(orc *o) {
  for (int I = 3; I <= o->MX - 2; ++I) {
    int jm = JMF(I, 0);
    QZF(I, 0) = rexp(-TAB(T_REL3_QZ, jm)) * QZF(I, 0) - TAB(T_REL1_QZ, jm) * DZ0(TZZF, I) -
                TAB(T_REL2_QZ, jm) * DZ0(PRF, I);
  }
}

/* INH_V_W_INT_0, inh_v_w_int_0.F90:43-63 */
static void inh_v_w_int_0(orc *o) {
  for (int I = 3; I <= o->MX - 2; ++I) {
    int jm = JMF(I, 0);
    WM(I, 0) = WM(I, 0) +
               (TAB(T_REL2_W, jm) / TAB(T_REL3_QZ, jm)) * ((real)1 - rexp(-TAB(T_REL3_QZ, jm))) *
                   QZPF(I, 0) +
               TAB(T_REL1_W, jm) * DZ0(TZZF, I) + TAB(T_REL3_W, jm) * DZ0(PRF, I);
  }
}

/* INH_QX_2ND_0, inh_qx_2nd_0.F90:55-104, and INH_V_U_2ND_0, inh_v_u_2nd_0.F90:57-106:
 * the side strips of row 0 (the z term is TXZ(i,1), TXZ(i,0) being 0) */
static void inh_qxu_2nd_0(orc *o) {
  const int L = 0;
  for (int side = 0; side < 2; ++side) {
    int lo = side ? RIGHT_LO : LEFT_UQX_LO, hi = side ? RIGHT_HI : LEFT_UQX_HI;
    for (int I = lo; I <= hi; ++I) {
      int jm = JMF(I, L);
      int j = side ? I - o->MX1 : 3 - I;
      real oda = side ? XL2(o->ODA2_XH, j, L) : XL2(o->ODA2_XL, j, L);
      real odb = side ? XL2(o->ODB2_XH, j, L) : XL2(o->ODB2_XL, j, L);
      real odc = side ? XL2(o->ODC2_XH, j, L) : XL2(o->ODC2_XL, j, L);
      real odd = side ? XL2(o->ODD2_XH, j, L) : XL2(o->ODD2_XL, j, L);
      real *G = side ? &MR1(o->GXXH, W_R1, I, L) : &ML(o->GXXL, W_L3, I, L);
      real *Hh = side ? &MR1(o->HXXH, W_R1, I, L) : &ML(o->HXXL, W_L3, I, L);
      real DXXXA = TXXF(I, L) - TXXF(I - 1, L);
      real PXXXA = PRF(I, L) - PRF(I - 1, L);
      QXF(I, L) = rexp(-TAB(T_REL3_QX, jm)) * QXF(I, L) -
                  TAB(T_REL1_QX, jm) * (oda * (odd * DXXXA + *G) + TXZF(I, 1)) -
                  TAB(T_REL2_QX, jm) * (oda * (odd * PXXXA + *Hh));
      *G = odb * *G + odc * DXXXA;
      *Hh = odb * *Hh + odc * PXXXA;
    }
  }
  for (int side = 0; side < 2; ++side) {
    int lo = side ? RIGHT_LO : LEFT_UQX_LO, hi = side ? RIGHT_HI : LEFT_UQX_HI;
    for (int I = lo; I <= hi; ++I) {
      int jm = JMF(I, L);
      int j = side ? I - o->MX1 : 3 - I;
      real oda = side ? XL2(o->ODA2_XH, j, L) : XL2(o->ODA2_XL, j, L);
      real odb = side ? XL2(o->ODB2_XH, j, L) : XL2(o->ODB2_XL, j, L);
      real odc = side ? XL2(o->ODC2_XH, j, L) : XL2(o->ODC2_XL, j, L);
      real odd = side ? XL2(o->ODD2_XH, j, L) : XL2(o->ODD2_XL, j, L);
      real *R = side ? &MR1(o->RXXH, W_R1, I, L) : &ML(o->RXXL, W_L3, I, L);
      real *S = side ? &MR1(o->SXXH, W_R1, I, L) : &ML(o->SXXL, W_L3, I, L);
      real DXXXA = TXXF(I, L) - TXXF(I - 1, L);
      real PXXXA = PRF(I, L) - PRF(I - 1, L);
      UM(I, L) = UM(I, L) +
                 (TAB(T_REL2_U, jm) / TAB(T_REL3_QX, jm)) * ((real)1 - rexp(-TAB(T_REL3_QX, jm))) *
                     QXPF(I, L) +
                 TAB(T_REL1_U, jm) * (oda * (odd * DXXXA + *R) + TXZF(I, 1)) +
                 TAB(T_REL3_U, jm) * (oda * (odd * PXXXA + *S));
      *R = odb * *R + odc * DXXXA;
      *S = odb * *S + odc * PXXXA;
    }
  }
}

/* INH_QZ_2ND_0, inh_qz_2nd_0.F90:48-80, and INH_V_W_2ND_0, inh_v_w_2nd_0.F90:50-84:
 * the side strips of row 0 (no x-derivative term, so no PML memory) */
static void inh_qzw_2nd_0(orc *o) {
  const int MXT1 = o->MX - 1;
  for (int side = 0; side < 2; ++side) {
    int lo = side ? MXT1 : 2 - o->T, hi = side ? MXT1 + o->T : 2;
    for (int I = lo; I <= hi; ++I) {
      int jm = JMF(I, 0);
      QZF(I, 0) = rexp(-TAB(T_REL3_QZ, jm)) * QZF(I, 0) -
                  TAB(T_REL1_QZ, jm) * ((real)3 * TZZF(I, 0) - (real)1 / (real)3 * TZZF(I, 1)) -
                  TAB(T_REL2_QZ, jm) * ((real)3 * PRF(I, 0) - (real)1 / (real)3 * PRF(I, 1));
    }
  }
  for (int side = 0; side < 2; ++side) {
    int lo = side ? MXT1 : 2 - o->T, hi = side ? MXT1 + o->T : 2;
    for (int I = lo; I <= hi; ++I) {
      int jm = JMF(I, 0);
      WM(I, 0) = WM(I, 0) +
                 (TAB(T_REL2_W, jm) / TAB(T_REL3_QZ, jm)) * ((real)1 - rexp(-TAB(T_REL3_QZ, jm))) *
                     QZPF(I, 0) +
                 TAB(T_REL1_W, jm) * ((real)3 * TZZF(I, 0) - (real)1 / (real)3 * TZZF(I, 1)) +
                 TAB(T_REL3_W, jm) * ((real)3 * PRF(I, 0) - (real)1 / (real)3 * PRF(I, 1));
    }
  }
}

/* HPLANE_0_VELOCITY without stress imaging, hplane_0_velocity.F90:62-179, then
 * STO_QXP/STO_QZP of row 0 (:181-195) */
static void hplane_0_velocity(orc *o) {
  inh_qx_int_0(o);
  inh_v_uv_int_0(o);
  inh_qz_int_0(o);
  inh_v_w_int_0(o);
  inh_qxu_2nd_0(o);
  inh_qzw_2nd_0(o);
  sto_qp(o, 0);
}

/* INH_V_QZ_INT_1, inh_v_qz_int_1.F90:49-79, and INH_V_W_INT_1, inh_v_w_int_1.F90:51-82
 * (4th-order branch: the reference guards its 2nd-order branch with SEDOND) */
static void inh_qzw_int_1(orc *o) {
  const real A = o->A, B = o->B;
  for (int I = 3; I <= o->MX - 2; ++I) {
    int jm = JMF(I, 1);
    QZF(I, 1) = rexp(-TAB(T_REL3_QZ, jm)) * QZF(I, 1) -
                TAB(T_REL1_QZ, jm) *
                    DZ1L(A * (TXZF(I + 2, 1) - TXZF(I - 1, 1)) + B * (TXZF(I + 1, 1) - TXZF(I, 1)), TZZF, I) -
                TAB(T_REL2_QZ, jm) * DZ1(PRF, I);
  }
  for (int I = 3; I <= o->MX - 2; ++I) {
    int jm = JMF(I, 1);
    WM(I, 1) = WM(I, 1) +
               (TAB(T_REL2_W, jm) / TAB(T_REL3_QZ, jm)) * ((real)1 - rexp(-TAB(T_REL3_QZ, jm))) *
                   QZPF(I, 1) +
               TAB(T_REL1_W, jm) *
                   DZ1L(A * (TXZF(I + 2, 1) - TXZF(I - 1, 1)) + B * (TXZF(I + 1, 1) - TXZF(I, 1)), TZZF, I) +
               TAB(T_REL3_W, jm) * DZ1(PRF, I);
  }
}

/* HPLANE_1_VELOCITY without stress imaging, hplane_1_velocity.F90:79-207, then
 * STO_QXP/STO_QZP of row 1 */
static void hplane_1_velocity(orc *o) {
  inh_qx_int(o, 1);
  inh_v_uv_int(o, 1);
  inh_qzw_int_1(o);
  inh_qx_2nd(o, 1);
  inh_v_u_2nd(o, 1);
  inh_qz_2nd(o, 1);
  inh_v_w_2nd(o, 1);
  sto_qp(o, 1);
}

/* Bottom-boundary store, time_loop.F90:133-155 */
static void store_bottom(orc *o) {
  const int NX = o->NX;
  memcpy(o->SZ_UP, o->SZ_U, sizeof(real) * 3 * NX);
  memcpy(o->SZ_WP, o->SZ_W, sizeof(real) * 3 * NX);
  memcpy(o->SZ_QXP, o->SZ_QX, sizeof(real) * 3 * NX);
  memcpy(o->SZ_QZP, o->SZ_QZ, sizeof(real) * 3 * NX);
  for (int I = o->PMLL; I <= o->PMLXH; ++I) {
    size_t k = (size_t)(I - o->PMLL) * 3;
    for (int t = 0; t < 3; ++t) { /* TIKMZ0, TIKMZ1, TIKMZ2 */
      int L = o->PMLZH - t;
      o->SZ_U[k + t] = UM(I, L);
      o->SZ_W[k + t] = WM(I, L);
      o->SZ_QX[k + t] = QXF(I, L);
      o->SZ_QZ[k + t] = QZF(I, L);
    }
  }
}

/* HPLANE_PMLZH0_VELOCITY, hplane_pmlzh0_velocity.F90:34-63 */
static void hplane_pmlzh0_velocity(orc *o) {
  const int LST = o->PMLZH, LST1 = LST - 1, LST2 = LST - 2;
  const real *aw = o->ABW, *ah = o->ABH;
  for (int I = 1 - o->T; I <= o->MX + o->T; ++I) {
    size_t k = (size_t)(I - o->PMLL) * 3;
#define NRCOMB(c, F, S, SP)                                                                   \
  (c[0] * F[IXF(o, I, LST1)] + c[1] * F[IXF(o, I, LST2)] + c[2] * S[k + 0] + c[3] * S[k + 1] + \
   c[4] * S[k + 2] + c[5] * SP[k + 0] + c[6] * SP[k + 1] + c[7] * SP[k + 2])
    UM(I, LST) = NRCOMB(ah, o->f[F_U], o->SZ_U, o->SZ_UP);
    WM(I, LST) = NRCOMB(aw, o->f[F_W], o->SZ_W, o->SZ_WP);
    QXF(I, LST) = NRCOMB(ah, o->f[F_QX], o->SZ_QX, o->SZ_QXP);
    QZF(I, LST) = NRCOMB(aw, o->f[F_QZ], o->SZ_QZ, o->SZ_QZP);
#undef NRCOMB
  }
}

/* Receiver samples after a step: STORE_SEISMOGRAMS (store_seismograms.F90:47-83)
 * and STORE_SEISMOGRAMS_PRESSURE (store_seismograms_pressure.F90:45-58). The
 * time-derivative columns are formed by the output writer. */
static void sample_receivers(orc *o, int itile) {
  if (!o->traces) return;
  real *out = o->traces + (size_t)(itile - 1) * o->MR * 6;
  for (int j = 0; j < o->MR; ++j) {
    int ir = o->IRECX[j], lr = o->LRECX[j], iz = o->IRECZ[j], lz = o->LRECZ[j];
    int okx = ir >= o->PMLL && ir <= o->PMLXH && lr >= 0 && lr <= o->PMLZH;
    int okz = iz >= o->PMLL && iz <= o->PMLXH && lz >= 0 && lz <= o->PMLZH;
    int okp = iz >= o->PMLL && iz <= o->PMLXH && lr >= 0 && lr <= o->PMLZH;
    real *r = out + (size_t)j * 6;
    r[0] = okx ? UM(ir, lr) : 0;
    r[1] = okz ? -WM(iz, lz) : 0;
    r[2] = okx ? QXF(ir, lr) : 0;
    r[3] = okz ? -QZF(iz, lz) : 0;
    r[4] = okp ? PRF(iz, lr) : 0;
    r[5] = okp ? -TXXF(iz, lr) : 0;
  }
}

/* One DO ITILE iteration, time_loop.F90:91-190 (serial build). */
/* Row temporaries of OpenMP thread `tid` (the pool grows to the team size). */
static rowtmp thread_rows(orc *o, int tid) {
  real *b = o->rowpool + (size_t)tid * 5 * o->NX;
  rowtmp t = {b, b + o->NX, b + 2 * (size_t)o->NX, b + 3 * (size_t)o->NX, b + 4 * (size_t)o->NX};
  return t;
}

/* One DO ITILE iteration (time_loop.F90:91-190). The interior rows of each
 * half-step (HPLANE_IC_*, rows 2..MZ'-2) read only the other half-step's
 * fields and update only their own row and their own rows of the PML
 * memories, so they are independent: they are shared out over OpenMP threads,
 * each with its own row temporaries. Every cell is still evaluated by the same
 * expression, so the result does not depend on the thread count. */
static void one_step(orc *o, int itile) {
  const int MZ = o->MZ, PMLZH1 = o->PMLZH - 1;
  rowtmp t = {o->EXX, o->EZZ, o->EXZ, o->QXX, o->QZZ};
  real S = o->SOURTF[(size_t)(itile - 1) * 2 + 0];
  real F = o->SOURTF[(size_t)(itile - 1) * 2 + 1];
  const int nth = omp_get_max_threads();
  if (nth > o->nrowpool) {
    free(o->rowpool);
    o->rowpool = (real *)zalloc(sizeof(real) * 5 * (size_t)o->NX * (size_t)nth);
    o->nrowpool = nth;
  }
  /* stress half-step; the source row is HPLANE_LIN_STRESS */
#pragma omp parallel
  {
    rowtmp tt = thread_rows(o, omp_get_thread_num());
#pragma omp for schedule(static)
    for (int L = 2; L <= MZ - 2; ++L) hplane_ic_stress(o, L, &tt, L == o->LIN_Z, S, F);
  }
  if (o->SI) {
    hplane_01_stress_si(o);
  } else {
    hplane_0_stress(o, &t);
    hplane_1_stress(o, &t);
  }
  hplane_mz1_stress(o, &t);
  for (int L = MZ; L <= PMLZH1; ++L) hplane_pml_stress(o, L, L - MZ + 1, &t);
  hplane_pmlzh0_stress(o, &t);
  /* bottom store */
  store_bottom(o);
  /* velocity half-step */
#pragma omp parallel for schedule(static)
  for (int L = 2; L <= MZ - 2; ++L) hplane_ic_velocity(o, L);
  if (o->SI) {
    hplane_1_velocity_si(o);
  } else {
    hplane_0_velocity(o);
    hplane_1_velocity(o);
  }
  for (int L = MZ - 1; L <= PMLZH1; ++L) hplane_pml_velocity(o, L, L - MZ + 1);
  hplane_pmlzh0_velocity(o);
  sample_receivers(o, itile);
}

/* ========================================================== PML coefficients */

/* CALCUL_VELOCITY, calcul_velocity.F90:16-36 */
static void calcul_velocity(real SIG_1, real P, real AINV, real MU, real REL1, real REL2,
                            real REL3, real *VPF, real *VPS, real *VS) {
  real AI[2][2] = {{REL1, -REL3}, {-REL3, REL2}};
  real Bm[2][2] = {{SIG_1, P}, {P, AINV}};
  real TT[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) TT[i][j] = AI[i][0] * Bm[0][j] + AI[i][1] * Bm[1][j];
  real DELTA = (TT[0][0] + TT[1][1]) * (TT[0][0] + TT[1][1]) +
               (real)4 * (TT[0][1] * TT[1][0] - TT[0][0] * TT[1][1]);
#if defined(ORACLE_DOUBLE)
  real sd = sqrt(DELTA);
  *VPF = sqrt(((TT[0][0] + TT[1][1]) + sd) / (real)2);
  *VPS = sqrt(((TT[0][0] + TT[1][1]) - sd) / (real)2);
  *VS = sqrt(REL1 * MU);
#else
  real sd = sqrtf(DELTA);
  *VPF = sqrtf(((TT[0][0] + TT[1][1]) + sd) / (real)2);
  *VPS = sqrtf(((TT[0][0] + TT[1][1]) - sd) / (real)2);
  *VS = sqrtf(REL1 * MU);
#endif
  if (MU == 0) *VS = *VPF;
}

static real rlog(real x) {
#if defined(ORACLE_DOUBLE)
  return log(x);
#else
  return logf(x);
#endif
}
static real rpow(real x, real y) {
#if defined(ORACLE_DOUBLE)
  return pow(x, y);
#else
  return powf(x, y);
#endif
}

/* One damping profile (ppml.F90:169-207): fills a (0:T) and a2 (0:T+1) column. */
static void ppml_profile(orc *o, real OMG, real VEL, real SF, real WC, real POW, real *a,
                         real *b, real *c, real *d, real *a2, real *b2, real *c2, real *d2,
                         size_t s1, size_t s2) {
  const int T = o->T;
  const real DT = o->DT;
  for (int j = 0; j <= T; ++j) { a[j * s1] = 1; b[j * s1] = 1; c[j * s1] = 0; }
  for (int j = 0; j <= T + 1; ++j) { a2[j * s2] = 1; b2[j * s2] = 1; c2[j * s2] = 0; }
  for (int J = 1; J <= T + 1; ++J) {
    real OMGPDT, AP, DENOM;
    if (J <= T) {
      OMGPDT = OMG * VEL * rpow((real)J, POW);
      AP = (real)1. + SF * OMGPDT * (real)0.075 / (real)0.91;
      DENOM = rexp(-OMGPDT / AP - WC * DT);
      a[J * s1] = (DENOM + (real)1.) / (real)2. / AP;
      b[J * s1] = DENOM;
      c[J * s1] = OMGPDT / (OMGPDT + WC * AP * DT) * (DENOM - (real)1.);
    }
    OMGPDT = OMG * VEL * rpow((real)J - (real)0.5, POW);
    AP = (real)1. + SF * OMGPDT * (real)0.075 / (real)0.91;
    DENOM = rexp(-OMGPDT / AP - WC * DT);
    a2[J * s2] = (DENOM + (real)1.) / (real)2. / AP;
    b2[J * s2] = DENOM;
    c2[J * s2] = OMGPDT / (OMGPDT + WC * AP * DT) * (DENOM - (real)1.);
  }
  for (int j = 0; j <= T; ++j) d[j * s1] = (c[j * s1] + (real)2.) / (b[j * s1] + (real)1.);
  for (int j = 0; j <= T + 1; ++j) d2[j * s2] = (c2[j * s2] + (real)2.) / (b2[j * s2] + (real)1.);
}

/* PPML, ppml.F90:159-306 (serial build) */
static void ppml(orc *o, real R, real TAU, real POW, real SF, real WC) {
  const int T = o->T;
  const real H = o->H, DT = o->DT;
  real OMG;
  if (T == 0)
    OMG = 0;
  else
    OMG = -TAU * rlog(R) * DT / (real)T / rpow((real)T, POW) / H;
  for (int L = 0; L <= o->PMLZH; ++L) {
    for (int side = 0; side < 2; ++side) {
      int jm = JMF(side ? o->PMLXH : o->PMLL, L);
      real VPF, VPS, VS;
      calcul_velocity(TAB(T_SXX1, jm), -TAB(T_P1, jm), -TAB(T_P3, jm), TAB(T_MU, jm),
                      H * TAB(T_REL1_U, jm) / DT, H * TAB(T_REL2_QX, jm) / DT,
                      H * TAB(T_REL3_U, jm) / DT, &VPF, &VPS, &VS);
      size_t b1 = (size_t)L * (T + 1), b2 = (size_t)L * (T + 2);
      if (!side)
        ppml_profile(o, OMG, VS, SF, WC, POW, o->ODA_XL + b1, o->ODB_XL + b1, o->ODC_XL + b1,
                     o->ODD_XL + b1, o->ODA2_XL + b2, o->ODB2_XL + b2, o->ODC2_XL + b2,
                     o->ODD2_XL + b2, 1, 1);
      else
        ppml_profile(o, OMG, VS, SF, WC, POW, o->ODA_XH + b1, o->ODB_XH + b1, o->ODC_XH + b1,
                     o->ODD_XH + b1, o->ODA2_XH + b2, o->ODB2_XH + b2, o->ODC2_XH + b2,
                     o->ODD2_XH + b2, 1, 1);
    }
  }
  if (o->KTBO != 0) OMG = 0;
  for (int I = o->PMLL; I <= o->PMLXH; ++I) {
    int jm = JMF(I, o->PMLZH);
    real VPF, VPS, VS;
    calcul_velocity(TAB(T_SZZ2, jm), -TAB(T_P2, jm), -TAB(T_P3, jm), TAB(T_MU, jm),
                    H * TAB(T_REL1_W, jm) / DT, H * TAB(T_REL2_QZ, jm) / DT,
                    H * TAB(T_REL3_W, jm) / DT, &VPF, &VPS, &VS);
    size_t b1 = (size_t)(I - o->PMLL) * (T + 1), b2 = (size_t)(I - o->PMLL) * (T + 2);
    ppml_profile(o, OMG, VPF, SF, WC, POW, o->ODA_ZH + b1, o->ODB_ZH + b1, o->ODC_ZH + b1,
                 o->ODD_ZH + b1, o->ODA2_ZH + b2, o->ODB2_ZH + b2, o->ODC2_ZH + b2,
                 o->ODD2_ZH + b2, 1, 1);
  }
}

/* ================================================================= C ABI */

static void oset_err(orc *o, const char *m) {
  if (o) snprintf(o->err, sizeof o->err, "%s", m);
}

EXPORT int oracle_real_size(void) { return (int)sizeof(real); }

EXPORT int oracle_create(const oracle_desc *d, orc **out) {
  if (!d || !out) return -1;
  *out = NULL;
  if (d->mx < 8 || d->mz < 6 || d->tpml < 1 || d->tpml > 256 || d->jmnum < 1) return -2;
  if (d->variant & ~(V_AFDA_ZZ | V_AFDA_XZ | V_SECOND)) return -3;
  if (d->lin_z < 2 || d->lin_z > d->mz - 2 || d->lin_x < 2 - d->tpml ||
      d->lin_x > d->mx - 1 + d->tpml)
    return -4;
  orc *o = (orc *)zalloc(sizeof(orc));
  o->MX = d->mx; o->MZ = d->mz; o->T = d->tpml; o->SI = d->stress_imaging != 0;
  o->AFDA_ZZ = (d->variant & V_AFDA_ZZ) != 0; o->AFDA_XZ = (d->variant & V_AFDA_XZ) != 0;
  o->SECOND = (d->variant & V_SECOND) != 0;
  o->LIN_X = d->lin_x; o->LIN_Z = d->lin_z; o->KTBO = d->ktbo; o->MT2 = d->mt2;
  o->JMNUM = d->jmnum; o->next_step = 1;
  o->H = (real)d->h; o->DT = (real)d->dt;
  /* mod_auxil.F90:19-20 and aux1.F90:36-46 */
  o->A = o->SECOND ? (real)0. : (real)-1. / (real)24.;
  o->B = o->SECOND ? (real)1. : (real)9. / (real)8.;
  o->MX1 = o->MX - 1;
  o->PMLL = 1 - o->T; o->PMLXH = o->MX + o->T; o->PMLZH = o->MZ + o->T;
  o->NX = o->PMLXH - o->PMLL + 1; o->NZ = o->PMLZH + 1;
  o->AH = o->A / o->H; o->BH = o->B / o->H; o->CH = (real)1. / o->H;
  size_t ncell = (size_t)o->NX * o->NZ;
  const int T = o->T, NX = o->NX, NZ = o->NZ;
  o->JM = (int32_t *)zalloc(sizeof(int32_t) * ncell);
  for (int k = 0; k < T_N; ++k) o->tab[k] = (real *)zalloc(sizeof(real) * o->JMNUM);
  for (int k = 0; k < F_NFIELDS; ++k) o->f[k] = (real *)zalloc(sizeof(real) * ncell);
  o->EXX = (real *)zalloc(sizeof(real) * NX); o->EZZ = (real *)zalloc(sizeof(real) * NX);
  o->EXZ = (real *)zalloc(sizeof(real) * NX); o->QXX = (real *)zalloc(sizeof(real) * NX);
  o->QZZ = (real *)zalloc(sizeof(real) * NX);
#define AZ(p, n) p = (real *)zalloc(sizeof(real) * (size_t)(n))
  AZ(o->ODA_XL, (T + 1) * NZ); AZ(o->ODB_XL, (T + 1) * NZ); AZ(o->ODC_XL, (T + 1) * NZ);
  AZ(o->ODD_XL, (T + 1) * NZ); AZ(o->ODA_XH, (T + 1) * NZ); AZ(o->ODB_XH, (T + 1) * NZ);
  AZ(o->ODC_XH, (T + 1) * NZ); AZ(o->ODD_XH, (T + 1) * NZ);
  AZ(o->ODA_ZH, (T + 1) * NX); AZ(o->ODB_ZH, (T + 1) * NX); AZ(o->ODC_ZH, (T + 1) * NX);
  AZ(o->ODD_ZH, (T + 1) * NX);
  AZ(o->ODA2_XL, (T + 2) * NZ); AZ(o->ODB2_XL, (T + 2) * NZ); AZ(o->ODC2_XL, (T + 2) * NZ);
  AZ(o->ODD2_XL, (T + 2) * NZ); AZ(o->ODA2_XH, (T + 2) * NZ); AZ(o->ODB2_XH, (T + 2) * NZ);
  AZ(o->ODC2_XH, (T + 2) * NZ); AZ(o->ODD2_XH, (T + 2) * NZ);
  AZ(o->ODA2_ZH, (T + 2) * NX); AZ(o->ODB2_ZH, (T + 2) * NX); AZ(o->ODC2_ZH, (T + 2) * NX);
  AZ(o->ODD2_ZH, (T + 2) * NX);
  AZ(o->PZXL, (T + 2) * NZ); AZ(o->PZXH, (T + 1) * NZ); AZ(o->PXXL, (T + 1) * NZ);
  AZ(o->PXXH, (T + 1) * NZ); AZ(o->OXXL, (T + 1) * NZ); AZ(o->OXXH, (T + 1) * NZ);
  AZ(o->PXZH, (T + 2) * NX); AZ(o->PZZH, (T + 1) * NX); AZ(o->OZZH, (T + 1) * NX);
  AZ(o->RXXL, (T + 3) * NZ); AZ(o->SXXL, (T + 3) * NZ); AZ(o->GXXL, (T + 3) * NZ);
  AZ(o->HXXL, (T + 3) * NZ);
  AZ(o->RXXH, (T + 2) * NZ); AZ(o->SXXH, (T + 2) * NZ); AZ(o->GXXH, (T + 2) * NZ);
  AZ(o->HXXH, (T + 2) * NZ);
  AZ(o->RZXL, (T + 2) * NZ); AZ(o->GZXL, (T + 2) * NZ); AZ(o->RZXH, (T + 2) * NZ);
  AZ(o->GZXH, (T + 2) * NZ);
  AZ(o->RXZH, (T + 1) * NX); AZ(o->GXZH, (T + 1) * NX); AZ(o->RZZH, (T + 1) * NX);
  AZ(o->SZZH, (T + 1) * NX); AZ(o->GZZH, (T + 1) * NX); AZ(o->HZZH, (T + 1) * NX);
  AZ(o->SZ_U, 3 * NX); AZ(o->SZ_W, 3 * NX); AZ(o->SZ_QX, 3 * NX); AZ(o->SZ_QZ, 3 * NX);
  AZ(o->SZ_UP, 3 * NX); AZ(o->SZ_WP, 3 * NX); AZ(o->SZ_QXP, 3 * NX); AZ(o->SZ_QZP, 3 * NX);
  AZ(o->SOURTF, 2 * (size_t)(o->MT2 > 0 ? o->MT2 : 1));
#undef AZ
  *out = o;
  return 0;
}

EXPORT void oracle_destroy(orc *o) {
  if (!o) return;
  free(o->JM);
  for (int k = 0; k < T_N; ++k) free(o->tab[k]);
  for (int k = 0; k < F_NFIELDS; ++k) free(o->f[k]);
  free(o->rowpool);
  real *arrs[] = {o->EXX, o->EZZ, o->EXZ, o->QXX, o->QZZ, o->ODA_XL, o->ODB_XL, o->ODC_XL,
                  o->ODD_XL, o->ODA_XH, o->ODB_XH, o->ODC_XH, o->ODD_XH, o->ODA_ZH, o->ODB_ZH,
                  o->ODC_ZH, o->ODD_ZH, o->ODA2_XL, o->ODB2_XL, o->ODC2_XL, o->ODD2_XL,
                  o->ODA2_XH, o->ODB2_XH, o->ODC2_XH, o->ODD2_XH, o->ODA2_ZH, o->ODB2_ZH,
                  o->ODC2_ZH, o->ODD2_ZH, o->PZXL, o->PZXH, o->PXXL, o->PXXH, o->OXXL,
                  o->OXXH, o->PXZH, o->PZZH, o->OZZH, o->RXXL, o->RXXH, o->SXXL, o->SXXH,
                  o->GXXL, o->GXXH, o->HXXL, o->HXXH, o->RZXL, o->RZXH, o->GZXL, o->GZXH,
                  o->RXZH, o->GXZH, o->RZZH, o->SZZH, o->GZZH, o->HZZH, o->SZ_U, o->SZ_W,
                  o->SZ_QX, o->SZ_QZ, o->SZ_UP, o->SZ_WP, o->SZ_QXP, o->SZ_QZP, o->SOURTF,
                  o->traces};
  for (size_t k = 0; k < sizeof arrs / sizeof arrs[0]; ++k) free(arrs[k]);
  free(o->IRECX); free(o->LRECX); free(o->IRECZ); free(o->LRECZ);
  free(o);
}

EXPORT const char *oracle_last_error(orc *o) { return o ? o->err : "null handle"; }

/* jm: [NZ][NX] int32 (1-based material ids, already replicated into the PML
 * and image rows as read_model.F90:59-87 does); tab: [JMNUM][24], DT/H-scaled
 * as read_model_par.F90:63-74 leaves it. */
EXPORT int oracle_upload_model(orc *o, const int32_t *jm, const real *tab) {
  size_t ncell = (size_t)o->NX * o->NZ;
  for (size_t k = 0; k < ncell; ++k) {
    if (jm[k] < 1 || jm[k] > o->JMNUM) { oset_err(o, "material index out of range"); return -5; }
    o->JM[k] = jm[k];
  }
  for (int m = 0; m < o->JMNUM; ++m)
    for (int k = 0; k < T_N; ++k) o->tab[k][m] = tab[(size_t)m * T_N + k];
  return 0;
}

EXPORT int oracle_build_pml(orc *o, double r, double tau, double pw, double sf, double wc) {
  ppml(o, (real)r, (real)tau, (real)pw, (real)sf, (real)wc);
  return 0;
}

/* abw/abh: ABW01,ABW02,ABW10,ABW11,ABW12,ABW20,ABW21,ABW22 (mod_nr_bound.F90:47-52) */
EXPORT int oracle_set_bottom_nr(orc *o, const real *abw, const real *abh) {
  for (int k = 0; k < 8; ++k) { o->ABW[k] = abw[k]; o->ABH[k] = abh[k]; }
  return 0;
}

/* sourtf: SOURTF_A(2,MT2), solid then fluid per step (open_sr_file.F90:78-110) */
EXPORT int oracle_set_source(orc *o, const real *sourtf) {
  memcpy(o->SOURTF, sourtf, sizeof(real) * 2 * (size_t)o->MT2);
  return 0;
}

EXPORT int oracle_set_receivers(orc *o, int mr, const int *irecx, const int *lrecx,
                                const int *irecz, const int *lrecz) {
  free(o->IRECX); free(o->LRECX); free(o->IRECZ); free(o->LRECZ); free(o->traces);
  o->MR = mr;
  o->IRECX = (int *)zalloc(sizeof(int) * mr); o->LRECX = (int *)zalloc(sizeof(int) * mr);
  o->IRECZ = (int *)zalloc(sizeof(int) * mr); o->LRECZ = (int *)zalloc(sizeof(int) * mr);
  memcpy(o->IRECX, irecx, sizeof(int) * mr); memcpy(o->LRECX, lrecx, sizeof(int) * mr);
  memcpy(o->IRECZ, irecz, sizeof(int) * mr); memcpy(o->LRECZ, lrecz, sizeof(int) * mr);
  o->traces = (real *)zalloc(sizeof(real) * (size_t)o->MT2 * mr * 6);
  return 0;
}

EXPORT int oracle_run(orc *o, int itile_begin, int itile_end) {
  if (itile_begin != o->next_step) { oset_err(o, "steps must run in order"); return -6; }
  if (itile_end > o->MT2) { oset_err(o, "step beyond MT2"); return -7; }
  for (int it = itile_begin; it <= itile_end; ++it) one_step(o, it);
  o->next_step = itile_end + 1;
  return 0;
}

EXPORT int oracle_get_field(orc *o, int id, real *host) {
  if (id < 0 || id >= F_NFIELDS) return -8;
  memcpy(host, o->f[id], sizeof(real) * (size_t)o->NX * o->NZ);
  return 0;
}

EXPORT int oracle_set_field(orc *o, int id, const real *host) {
  if (id < 0 || id >= F_NFIELDS) return -8;
  memcpy(o->f[id], host, sizeof(real) * (size_t)o->NX * o->NZ);
  return 0;
}

EXPORT int oracle_get_traces(orc *o, real *out, int first_step, int nsteps) {
  if (!o->traces || first_step < 1 || first_step + nsteps - 1 > o->MT2) return -9;
  memcpy(out, o->traces + (size_t)(first_step - 1) * o->MR * 6,
         sizeof(real) * (size_t)nsteps * o->MR * 6);
  return 0;
}

/* Copies one PML coefficient array out: id 0..23 in mod_pml.F90 order
 * (ODA_XL ODB_XL ODC_XL ODA_XH ODB_XH ODC_XH ODD_XL ODD_XH ODA_ZH ODB_ZH ODC_ZH
 *  ODD_ZH ODA2_XL ODB2_XL ODC2_XL ODA2_XH ODB2_XH ODC2_XH ODD2_XL ODD2_XH
 *  ODA2_ZH ODB2_ZH ODC2_ZH ODD2_ZH). Returns the element count. */
EXPORT int oracle_get_pml(orc *o, int id, real *out) {
  real *tbl[24] = {o->ODA_XL, o->ODB_XL, o->ODC_XL, o->ODA_XH, o->ODB_XH, o->ODC_XH,
                   o->ODD_XL, o->ODD_XH, o->ODA_ZH, o->ODB_ZH, o->ODC_ZH, o->ODD_ZH,
                   o->ODA2_XL, o->ODB2_XL, o->ODC2_XL, o->ODA2_XH, o->ODB2_XH, o->ODC2_XH,
                   o->ODD2_XL, o->ODD2_XH, o->ODA2_ZH, o->ODB2_ZH, o->ODC2_ZH, o->ODD2_ZH};
  if (id < 0 || id >= 24) return -8;
  size_t n;
  if (id < 8) n = (size_t)(o->T + 1) * o->NZ;
  else if (id < 12) n = (size_t)(o->T + 1) * o->NX;
  else if (id < 20) n = (size_t)(o->T + 2) * o->NZ;
  else n = (size_t)(o->T + 2) * o->NX;
  if (out) memcpy(out, tbl[id], sizeof(real) * n);
  return (int)n;
}

/* ====================================================== source time function */

/* locate, open_sr_file.F90:200-226 (1-based result) */
static int sp_locate(const real *xx, int n, real x) {
  int ascnd = xx[n - 1] >= xx[0];
  int jl = 0, ju = n + 1;
  while (ju - jl > 1) {
    int jm = (ju + jl) / 2;
    if (ascnd == (x >= xx[jm - 1])) jl = jm; else ju = jm;
  }
  if (x == xx[0]) return 1;
  if (x == xx[n - 1]) return n - 1;
  return jl;
}

/* spline + tridag, open_sr_file.F90:121-175, with yp1 = ypn = 0 (:81) */
static int sp_spline(const real *x, const real *y, int n, real *y2) {
  real *a = (real *)zalloc(sizeof(real) * n), *b = (real *)zalloc(sizeof(real) * n);
  real *c = (real *)zalloc(sizeof(real) * n), *r = (real *)zalloc(sizeof(real) * n);
  real *gam = (real *)zalloc(sizeof(real) * n), *r0 = (real *)zalloc(sizeof(real) * n);
  const real yp1 = 0, ypn = 0;
  for (int i = 0; i < n - 1; ++i) c[i] = x[i + 1] - x[i];
  for (int i = 0; i < n - 1; ++i) r[i] = (real)6.0 * ((y[i + 1] - y[i]) / c[i]);
  for (int i = 0; i < n - 1; ++i) r0[i] = r[i];
  for (int i = 1; i < n - 1; ++i) r[i] = r0[i] - r0[i - 1];
  for (int i = 1; i < n - 1; ++i) a[i] = c[i - 1];
  for (int i = 1; i < n - 1; ++i) b[i] = (real)2.0 * (c[i] + a[i]);
  b[0] = 1; b[n - 1] = 1;
  r[0] = ((real)3.0 / (x[1] - x[0])) * ((y[1] - y[0]) / (x[1] - x[0]) - yp1);
  c[0] = (real)0.5;
  r[n - 1] = ((real)-3.0 / (x[n - 1] - x[n - 2])) * ((y[n - 1] - y[n - 2]) / (x[n - 1] - x[n - 2]) - ypn);
  a[n - 1] = (real)0.5;
  /* tridag(a(2:n), b(1:n), c(1:n-1), r(1:n), y2): its a(j-1) is a[j-1+1-1] */
  int rc = 0;
  real bet = b[0];
  if (bet == 0) rc = -1;
  y2[0] = r[0] / bet;
  for (int j = 1; j < n && !rc; ++j) {
    gam[j] = c[j - 1] / bet;
    bet = b[j] - a[j] * gam[j];
    if (bet == 0) { rc = -1; break; }
    y2[j] = (r[j] - a[j] * y2[j - 1]) / bet;
  }
  for (int j = n - 2; j >= 0 && !rc; --j) y2[j] = y2[j] - gam[j + 1] * y2[j + 1];
  free(a); free(b); free(c); free(r); free(gam); free(r0);
  return rc;
}

/* splint, open_sr_file.F90:177-198 */
static real sp_splint(const real *xa, const real *ya, const real *y2a, int n, real x) {
  int klo = imax(imin(sp_locate(xa, n, x), n - 1), 1);
  int khi = klo + 1;
  real h = xa[khi - 1] - xa[klo - 1];
  real a = (xa[khi - 1] - x) / h;
  real b = (x - xa[klo - 1]) / h;
  return a * ya[klo - 1] + b * ya[khi - 1] +
         ((a * a * a - a) * y2a[klo - 1] + (b * b * b - b) * y2a[khi - 1]) * (h * h) / (real)6.0;
}

/* OPEN_SR_FILE body, open_sr_file.F90:62-110: nstf samples (t, v) as read from
 * STF.DAT -> SOURTF_A(2,MT2) (solid, fluid per step). Returns 0 or < 0. */
EXPORT int oracle_source_from_stf(const real *t, const real *v, int nstf, int mt2, double dt_,
                                  double h_, int key_sour, double qr_, real *sourtf) {
  if (nstf < 1 || mt2 < 1) return -1;
  const real DT = (real)dt_, H = (real)h_, QR = (real)qr_;
  int n = nstf + 1;
  real *x = (real *)zalloc(sizeof(real) * n), *y = (real *)zalloc(sizeof(real) * n);
  real *y2 = (real *)zalloc(sizeof(real) * n);
  memcpy(x, t, sizeof(real) * nstf);
  memcpy(y, v, sizeof(real) * nstf);
  x[n - 1] = (real)mt2 * DT;
  y[n - 1] = y[n - 2];
  int rc = sp_spline(x, y, n, y2);
  for (int k = 0; k < 2 * mt2; ++k) sourtf[k] = 0;
  for (int I = 1; I <= mt2 && !rc; ++I) {
    real T = (real)(I - 1) * DT;
    if (T > x[n - 2]) break;
    real s = sp_splint(x, y, y2, n, T);
    real *o = sourtf + (size_t)(I - 1) * 2;
    switch (key_sour) {
      case 0: o[1] = QR * s / (H * H); o[0] = ((real)1 - QR) * s / (H * H); break;
      case 1: o[1] = 0; o[0] = s / (H * H); break;
      case 2: o[1] = s / (H * H); o[0] = QR * s / (H * H); break;
      case 3: o[1] = s / (H * H); o[0] = 0; break;
      default: o[0] = s; break;
    }
  }
  free(x); free(y); free(y2);
  return rc;
}

/* ====================================================== bottom boundary keys */

/* PRE_NR_COEF1 bottom part (pre_nr_coef1.F90:185-201) with NONREF
 * (nonref.F90:69-204) for the closed-form keys 0, 2, 4 and 5: wave speeds of
 * the cell JM(PMLL, MZ'), ABW from the fast-P speed, ABH from the S speed.
 * Returns -1 for keys whose coefficients come from HIGDON_FD/OptAbsBc. */
/* Parameters of MODULE CONTROL_DATA that NONREF reads: OMG as READ_OPEN leaves
 * it (namelist value / 2 pi, read_open.F90:188), WB, THPPLE, THSSLE
 * (read_open.F90:173-179 defaults). */
typedef struct { real OMG, WB, THPPLE, THSSLE; } nr_params;

/* HIGDON_FD, higdon_fd.F90:1-46 (DP arithmetic on PP arguments) */
static void higdon_fd(real WB, real ALFA, real BETA, real DT, real H, double *a) {
  double BE1 = 1.0;
  double BE2 = (double)ALFA / (double)BETA;
  double NI = (double)ALFA * (double)DT / (double)H;
  double QX = (WB * (BE1 + NI) - NI) / (BE1 + NI) / (1.0 - WB);
  double RX = (WB * (BE2 + NI) - NI) / (BE2 + NI) / (1.0 - WB);
  double QT = (WB * (BE1 + NI) - BE1) / (BE1 + NI) / (1.0 - WB);
  double RT = (WB * (BE2 + NI) - BE1) / (BE2 + NI) / (1.0 - WB);
  double QXT = WB / (WB - 1.0);
  double RXT = WB / (WB - 1.0);
  a[0] = 0.0;                                  /* A00 */
  a[1] = -(QX + RX);                           /* A01 */
  a[2] = -(QX * RX);                           /* A02 */
  a[3] = -(QT + RT);                           /* A10 */
  a[4] = -(QX * RT + QT * RX + QXT + RXT);     /* A11 */
  a[5] = -(QX * RXT + RX * QXT);               /* A12 */
  a[6] = -(QT * RT);                           /* A20 */
  a[7] = -(QT * RXT + RT * QXT);               /* A21 */
  a[8] = -(QXT * RXT);                         /* A22 */
}

/* HIGCE_FD, higdce_fd.F90:1-42: GP, GS are formed in PP, and so is (1-WB) */
static void higce_fd(real WB, real ALFA, real BETA, real DT, real H, double *a) {
  double GP = (real)(ALFA * DT / H);
  double GS = (real)(BETA * DT / H);
  double GS1 = 1.0 + GS;
  real OMW = (real)1 - WB;
  double H1X = (GS - WB * GS1) / (GS1 * OMW);
  double H1T = (1.0 - WB * GS1) / (GS1 * OMW);
  double HXT = WB / (1.0 - WB);
  double C1T = 1.0 - GP;
  double CXT = GP;
  a[0] = 0.0;
  a[1] = H1X;
  a[2] = 0.0;
  a[3] = C1T + H1T;
  a[4] = CXT + HXT - H1X * C1T;
  a[5] = -(H1X * CXT);
  a[6] = -(H1T * C1T);
  a[7] = -(HXT * C1T) - H1T * CXT;
  a[8] = -(HXT * CXT);
}

/* HIGCE2_FD, higdce2_fd.F90:1-45 */
static void higce2_fd(real WB, real VELH, real VELCE, real DT, real H, double *a) {
  double NI = (double)VELH * (double)DT / (double)H;
  double G = (double)VELCE * (double)DT / (double)H;
  double QX = (WB * (1.0 + NI) - NI) / (1.0 + NI) / (1.0 - WB);
  double RX = (1.0 - G) / (1.0 + G);
  double QT = (WB * (1.0 + NI) - 1.0) / (1.0 + NI) / (1.0 - WB);
  double RT = -((1.0 - G) / (1.0 + G));
  double QXT = WB / (WB - 1.0);
  double RXT = -1.0;
  a[0] = 0.0;
  a[1] = -(QX + RX);
  a[2] = -(QX * RX);
  a[3] = -(QT + RT);
  a[4] = -(QX * RT + QT * RX + QXT + RXT);
  a[5] = -(QX * RXT + RX * QXT);
  a[6] = -(QT * RT);
  a[7] = -(QT * RXT + RT * QXT);
  a[8] = -(QXT * RXT);
}

/* OptAbsBc, OptAbsBc.F90:1-106 (DP; the literals are exact binary values) */
static void opt_abs_bc(double gplus, double gminus, double dt, double dx, double *a) {
  double tplus = dx * cos(gplus);
  double tminus = dx * cos(gminus);
  double Rz = cos(dt), Iz = sin(dt);
  double Rx = cos(tplus + tminus), Ix = sin(tplus + tminus);
  double Ry = -(cos(tplus) + cos(tminus));
  double Iy = -(sin(tplus) + sin(tminus));
  double temp1 = Rx * Rz - Ix * Iz;
  double temp2 = Rx * Iz + Rz * Ix;
  double temp3 = Ry * Rz - Iy * Iz;
  double temp4 = Ry * Iz + Rz * Iy;
  double delta = (Ry + Iy + 2.0 * Rz) * (temp1 + temp2 - Rz + Iz) +
                 (1.0 - Rx - Ix) * (2.0 + temp3 + temp4);
  delta = 1.0 / delta;
  double r1[3][3], r0[3][3];
  r1[1][0] = (1.0 - Rz) * (temp1 + temp2 - Rz + Iz) * delta;
  r1[0][2] = (1.0 - Rz) * (1.0 - Rx - Ix) * delta;
  r1[0][1] = -0.50 - r1[1][0];
  r1[1][1] = 1.0 - 2.0 * r1[0][2];
  r1[0][0] = 0.0;
  r1[2][2] = 0.0;
  r1[1][2] = r1[1][0];
  r1[2][0] = r1[0][2];
  r1[2][1] = r1[0][1];
  double temp5 = Rx * Rz + Ix * Iz;
  double temp6 = Rx * Iz - Rz * Ix;
  double temp7 = Ry * Rz + Iy * Iz;
  double temp8 = Ry * Iz - Rz * Iy;
  double phi = -((Rz + Iz - temp5 + temp6) * (Ry + Iy + 2.0 * Rz)) +
               (1.0 + Rz + temp7 - temp8) * (1.0 - Rx - Ix);
  r0[0][2] = phi * delta;
  double psi = (1.0 + Rz + temp7 - temp8) * (temp1 + temp2 - Rz + Iz) +
               (Rz + Iz - temp5 + temp6) * (2.0 + temp3 + temp4);
  r0[1][0] = psi * delta;
  r0[0][0] = 0.;
  r0[0][1] = 0.50 - r0[1][0];
  r0[1][1] = 1.0 - 2.0 * r0[0][2];
  r0[2][2] = -1.0;
  r0[1][2] = r0[1][0];
  r0[2][0] = r0[0][2];
  r0[2][1] = r0[0][1];
  temp1 = 0.0;
  temp2 = 0.0;
  for (int i = 0; i <= 2; ++i)
    for (int j = 0; j <= 2; ++j) {
      temp1 = temp1 + r0[i][j] * r1[i][j];
      temp2 = temp2 + r1[i][j] * r1[i][j];
    }
  double ratio = -(temp1 / temp2);
  a[0] = 0.0;
  a[1] = r0[0][1] + ratio * r1[0][1];
  a[2] = r0[0][2] + ratio * r1[0][2];
  a[3] = r0[1][0] + ratio * r1[1][0];
  a[4] = r0[1][1] + ratio * r1[1][1];
  a[5] = r0[1][2] + ratio * r1[1][2];
  a[6] = r0[2][0] + ratio * r1[2][0];
  a[7] = r0[2][1] + ratio * r1[2][1];
  a[8] = r0[2][2] + ratio * r1[2][2];
}

/* NONREF, nonref.F90:56-206: ax = AX01, AX02, AX10, AX11, AX12, AX20, AX21, AX22.
 * Keys 3 and 9 read THPPLE and THSSLE whichever boundary is served, and hand
 * back REAL(aa) - default (binary32) reals. */
static int nonref_key(int key, real H, real DT, real VEL, real VELS, real VELP,
                      const nr_params *p, real *ax) {
  double aa[9];
  for (int k = 0; k < 8; ++k) ax[k] = 0;
  switch (key) {
    case 0: return 0;
    case 1: higdon_fd(p->WB, VELP, VELS, DT, H, aa); break;
    case 2: {
      real TMP = -(DT * VEL / H);
      ax[2] = (real)1 + TMP; ax[3] = (real)1 - TMP; ax[6] = (real)-1 - TMP; ax[7] = TMP;
      return 0;
    }
    case 3:
    case 9: {
      real HX = key == 3 ? p->OMG / VEL * H : p->OMG / VELS * H;
      opt_abs_bc((double)p->THPPLE, (double)p->THSSLE, (double)(real)(p->OMG * DT), (double)HX, aa);
      for (int k = 0; k < 8; ++k) ax[k] = (real)(float)aa[k + 1];
      return 0;
    }
    case 4: {
      real TMP1 = (DT - H / VEL) / (DT + H / VEL);
      real TMP2 = (real)2 * H / VEL / (DT + H / VEL);
      ax[0] = TMP1; ax[2] = TMP2; ax[3] = TMP2; ax[5] = TMP1; ax[6] = (real)-1;
      return 0;
    }
    case 5: {
      real TMP1 = VEL * DT / H;
      ax[0] = -((real)1 - TMP1) / ((real)1 + TMP1);
      ax[2] = ((real)1 - TMP1) / ((real)1 + TMP1);
      ax[3] = (real)1;
      return 0;
    }
    case 6: higce_fd(p->WB, VELP, VELS, DT, H, aa); break;
    case 7: higce2_fd(p->WB, VEL, VEL, DT, H, aa); break;
    case 8: higce2_fd(p->WB, VELS, VELP, DT, H, aa); break;
    default: return -1;
  }
  for (int k = 0; k < 8; ++k) ax[k] = (real)aa[k + 1];
  return 0;
}

/* PRE_NR_COEF1 bottom part, pre_nr_coef1.F90:185-201: ABW from the fast P
 * speed, ABH from the S speed of the material at (X_MIN, MZ'). prm = OMG (as
 * READ_OPEN leaves it), WB, THPPLE, THSSLE. */
EXPORT int oracle_bottom_nr(orc *o, int key, const real *prm, real *abw, real *abh) {
  int jm = JMF(o->PMLL, o->MZ);
  real VPF, VPS, VS;
  calcul_velocity(TAB(T_SXX1, jm), -TAB(T_P1, jm), -TAB(T_P3, jm), TAB(T_MU, jm),
                  o->H * TAB(T_REL1_U, jm) / o->DT, o->H * TAB(T_REL2_QX, jm) / o->DT,
                  o->H * TAB(T_REL3_U, jm) / o->DT, &VPF, &VPS, &VS);
  nr_params p = {prm[0], prm[1], prm[2], prm[3]};
  if (nonref_key(key, o->H, o->DT, VPF, VS, VPF, &p, abw)) return -1;
  if (nonref_key(key, o->H, o->DT, VS, VS, VPF, &p, abh)) return -1;
  return 0;
}

/* NONREF alone, for tests of the library's restatement. */
EXPORT int oracle_nonref(int key, real h, real dt, real vel, real vels, real velp, const real *prm,
                         real *ax) {
  nr_params p = {prm[0], prm[1], prm[2], prm[3]};
  return nonref_key(key, h, dt, vel, vels, velp, &p, ax);
}
