!=======================================================================
! TIME_LOOP replacement that runs the step loop on the MI355X.
!
! Drop-in for SUBROUTINE TIME_LOOP (time_loop.F90:1-209): same name, no
! arguments, same module state. The set-up the reference driver already did
! (READ_MODEL, READ_MODEL_PAR, OPEN_SR_FILE, PRE_NR_COEF1, PPML) is handed to
! the device once; the steps run there. Receiver samples come back once per
! chunk of steps. STORE_CHUNK below appends the chunk's lines to the receiver
! files with one OPEN per file per chunk: the same values and the same format
! STORE_SEISMOGRAMS and STORE_SEISMOGRAMS_PRESSURE write with one OPEN per file
! per step (store_seismograms.F90:54-153, store_seismograms_pressure.F90:45-81).
! Built with -DBQ_REFERENCE_WRITERS, the reference's own writers run instead,
! step by step, after the samples are placed at the receiver nodes of the host
! arrays. STORE_SNAPSHOTS writes the snapshots unchanged.
! The STOP file (time_loop.F90:79-86) is polled at the start of each chunk.
!=======================================================================
SUBROUTINE  TIME_LOOP

  USE ISO_C_BINDING
  USE BQ_BINDING
  USE PRECISION   , ONLY: PP
  USE CONTROL_DATA, ONLY: KEY_SNV, KEY_TLV, KEY_TLP, KEY_TLQ, MX, MZ, MT1, MT2, DT, IPAS2, &
                          TPML, LIN_X, LIN_Z, STRESS_IMAGING, KTBO, MR,                   &
                          IRECX, LRECX, IRECZ, LRECZ
  USE GRID_MEDIUM , ONLY: JM, JMNUM, H, MU,                                             &
                          SIG_XX_1, SIG_XX_2, SIG_XX_3, SIG_ZZ_1, SIG_ZZ_2, SIG_ZZ_3,   &
                          PRES_1, PRES_2, PRES_3, REL1_U, REL2_U, REL3_U,               &
                          REL1_W, REL2_W, REL3_W, REL1_QX, REL2_QX, REL3_QX,            &
                          REL1_QZ, REL2_QZ, REL3_QZ, AUXIL_W, AUXIL_QZ
  USE T_LOOP      , ONLY: ITILE
  USE AUXIL       , ONLY: SOURTF_A, X_MIN, X_MAX, Z_MIN, Z_MAX
  USE PML         , ONLY: ODA_XL, ODB_XL, ODC_XL, ODA_XH, ODB_XH, ODC_XH, ODD_XL, ODD_XH, &
                          ODA_ZH, ODB_ZH, ODC_ZH, ODD_ZH,                                 &
                          ODA2_XL, ODB2_XL, ODC2_XL, ODA2_XH, ODB2_XH, ODC2_XH,           &
                          ODD2_XL, ODD2_XH, ODA2_ZH, ODB2_ZH, ODC2_ZH, ODD2_ZH
  USE NONREF_BOUND, ONLY: ABW01, ABW02, ABW10, ABW11, ABW12, ABW20, ABW21, ABW22,         &
                          ABH01, ABH02, ABH10, ABH11, ABH12, ABH20, ABH21, ABH22
  USE WAVEFIELD   , ONLY: UM, WM, QX, QZ, SSU, SSW, SSQX, SSQZ
  USE STRESSFIELD , ONLY: TXX, PRES
  USE INTERFACES  , ONLY: STORE_SEISMOGRAMS, STORE_SEISMOGRAMS_PRESSURE, STORE_SNAPSHOTS

  IMPLICIT NONE

  INTERFACE
    SUBROUTINE     TL_ALLOC
    END SUBROUTINE TL_ALLOC
  END INTERFACE

  INTEGER, PARAMETER :: CHUNK_MAX = 256

  TYPE(BQ_DESC) :: D
  TYPE(C_PTR)   :: HB
  TYPE(C_PTR)   :: COEF(24)
  REAL(C_FLOAT), ALLOCATABLE, TARGET :: PBUF(:)
  REAL(C_FLOAT), ALLOCATABLE :: TAB(:,:), TR(:,:,:)
  REAL(C_FLOAT) :: ABW(8), ABH(8)
  INTEGER       :: RC, I0, I1, J, S, K, OFF, N
  LOGICAL       :: EX, SEIS
  CHARACTER(LEN=20), EXTERNAL :: STR

  CALL TL_ALLOC

  D%MX = MX
  D%MZ = MZ
  D%TPML = TPML
  D%STRESS_IMAGING = MERGE(1, 0, STRESS_IMAGING)
  D%LIN_X = LIN_X
  D%LIN_Z = LIN_Z
  D%KTBO = KTBO
  D%MT2 = MT2
  D%JMNUM = JMNUM
  ! the build options this driver was configured with (the reference's
  ! -DAFDA_ZZ / -DAFDA_XZ one-sided surface stencils, -DSECOND second order)
  D%VARIANT = 0
#if(AFDA_ZZ)
  D%VARIANT = IOR(D%VARIANT, BQ_AFDA_ZZ)
#endif
#if(AFDA_XZ)
  D%VARIANT = IOR(D%VARIANT, BQ_AFDA_XZ)
#endif
#if(SECOND)
  D%VARIANT = IOR(D%VARIANT, BQ_SECOND)
#endif
  D%H = REAL(H, C_DOUBLE)
  D%DT = REAL(DT, C_DOUBLE)
  RC = BQ_CREATE(D, HB)
  IF (RC /= 0) CALL FAIL('BQ_CREATE', RC)

  ! material table in the .MO order, as READ_MODEL_PAR left it (scaled)
  ALLOCATE (TAB(24, JMNUM))
  TAB( 1,:) = MU;       TAB( 2,:) = SIG_XX_1; TAB( 3,:) = SIG_XX_2; TAB( 4,:) = SIG_XX_3
  TAB( 5,:) = SIG_ZZ_1; TAB( 6,:) = SIG_ZZ_2; TAB( 7,:) = SIG_ZZ_3
  TAB( 8,:) = PRES_1;   TAB( 9,:) = PRES_2;   TAB(10,:) = PRES_3
  TAB(11,:) = REL1_U;   TAB(12,:) = REL1_W;   TAB(13,:) = REL2_U
  TAB(14,:) = REL2_W;   TAB(15,:) = REL3_U;   TAB(16,:) = REL3_W
  TAB(17,:) = REL1_QX;  TAB(18,:) = REL1_QZ;  TAB(19,:) = REL2_QX
  TAB(20,:) = REL2_QZ;  TAB(21,:) = REL3_QX;  TAB(22,:) = REL3_QZ
  TAB(23,:) = AUXIL_W;  TAB(24,:) = AUXIL_QZ
  RC = BQ_UPLOAD_MODEL(HB, JM, TAB)
  IF (RC /= 0) CALL FAIL('BQ_UPLOAD_MODEL', RC)
  DEALLOCATE (TAB)

  ! the 24 PPML arrays, in MODULE PML order, packed into one buffer
  N = 8*SIZE(ODA_XL) + 4*SIZE(ODA_ZH) + 8*SIZE(ODA2_XL) + 4*SIZE(ODA2_ZH)
  ALLOCATE (PBUF(N))
  OFF = 1
  CALL PUT(ODA_XL, 1);  CALL PUT(ODB_XL, 2);  CALL PUT(ODC_XL, 3)
  CALL PUT(ODA_XH, 4);  CALL PUT(ODB_XH, 5);  CALL PUT(ODC_XH, 6)
  CALL PUT(ODD_XL, 7);  CALL PUT(ODD_XH, 8)
  CALL PUT(ODA_ZH, 9);  CALL PUT(ODB_ZH, 10); CALL PUT(ODC_ZH, 11); CALL PUT(ODD_ZH, 12)
  CALL PUT(ODA2_XL, 13); CALL PUT(ODB2_XL, 14); CALL PUT(ODC2_XL, 15)
  CALL PUT(ODA2_XH, 16); CALL PUT(ODB2_XH, 17); CALL PUT(ODC2_XH, 18)
  CALL PUT(ODD2_XL, 19); CALL PUT(ODD2_XH, 20)
  CALL PUT(ODA2_ZH, 21); CALL PUT(ODB2_ZH, 22); CALL PUT(ODC2_ZH, 23); CALL PUT(ODD2_ZH, 24)
  RC = BQ_UPLOAD_PML(HB, COEF)
  IF (RC /= 0) CALL FAIL('BQ_UPLOAD_PML', RC)
  DEALLOCATE (PBUF)

  ABW = [ABW01, ABW02, ABW10, ABW11, ABW12, ABW20, ABW21, ABW22]
  ABH = [ABH01, ABH02, ABH10, ABH11, ABH12, ABH20, ABH21, ABH22]
  RC = BQ_UPLOAD_BOTTOM_NR(HB, ABW, ABH)
  IF (RC /= 0) CALL FAIL('BQ_UPLOAD_BOTTOM_NR', RC)

  RC = BQ_SET_SOURCE(HB, SOURTF_A)
  IF (RC /= 0) CALL FAIL('BQ_SET_SOURCE', RC)

  SEIS = (KEY_TLV .OR. KEY_TLP) .AND. MR > 0
  IF (SEIS) THEN
    RC = BQ_SET_RECEIVERS(HB, MR, IRECX, LRECX, IRECZ, LRECZ)
    IF (RC /= 0) CALL FAIL('BQ_SET_RECEIVERS', RC)
    ALLOCATE (TR(6, MR, CHUNK_MAX))
  END IF

  I0 = MT1
  DO WHILE (I0 <= MT2)
    INQUIRE ( FILE = 'STOP', EXIST = EX )
    IF ( EX ) THEN
      OPEN  ( 35, FILE='STOP', STATUS='UNKNOWN')
      WRITE ( 35,* ) I0
      CLOSE ( 35 )
      CALL BQ_DESTROY(HB)
      STOP
    END IF
    I1 = MIN(MT2, I0 + CHUNK_MAX - 1)
    IF (KEY_SNV) I1 = MIN(I1, ((I0 + IPAS2 - 1) / IPAS2) * IPAS2)
    RC = BQ_RUN(HB, I0, I1)
    IF (RC /= 0) CALL FAIL('BQ_RUN', RC)
    IF (SEIS) THEN
      RC = BQ_GET_TRACES(HB, TR, I0, I1 - I0 + 1)
      IF (RC /= 0) CALL FAIL('BQ_GET_TRACES', RC)
    ELSE
      RC = BQ_SYNC(HB)
      IF (RC /= 0) CALL FAIL('BQ_SYNC', RC)
    END IF
#if(!BQ_REFERENCE_WRITERS)
    DO ITILE = I0, I1
      WRITE (11,'(I6)') ITILE
    END DO
    IF (SEIS) CALL STORE_CHUNK(I0, I1)
#else
    DO ITILE = I0, I1
      WRITE (11,'(I6)') ITILE
      IF (SEIS) THEN
        S = ITILE - I0 + 1
        DO J = 1, MR
          IF (INSIDE(IRECX(J), LRECX(J))) THEN
            UM(IRECX(J), LRECX(J)) =  TR(1, J, S)
            QX(IRECX(J), LRECX(J)) =  TR(3, J, S)
          END IF
          IF (INSIDE(IRECZ(J), LRECZ(J))) THEN
            WM(IRECZ(J), LRECZ(J)) = -TR(2, J, S)
            QZ(IRECZ(J), LRECZ(J)) = -TR(4, J, S)
          END IF
          IF (INSIDE(IRECZ(J), LRECX(J))) THEN
            PRES(IRECZ(J), LRECX(J)) =  TR(5, J, S)
            TXX (IRECZ(J), LRECX(J)) = -TR(6, J, S)
          END IF
        END DO
        IF (KEY_TLV) CALL STORE_SEISMOGRAMS(ITILE)
        IF (KEY_TLP) CALL STORE_SEISMOGRAMS_PRESSURE(ITILE)
      END IF
    END DO
#endif
    IF (KEY_SNV .AND. MOD(I1, IPAS2) == 0) THEN
      RC = BQ_GET_FIELD(HB, BQ_U, UM)
      IF (RC == 0) RC = BQ_GET_FIELD(HB, BQ_W, WM)
      IF (RC == 0 .AND. KEY_TLQ) RC = BQ_GET_FIELD(HB, BQ_QX, QX)
      IF (RC == 0 .AND. KEY_TLQ) RC = BQ_GET_FIELD(HB, BQ_QZ, QZ)
      IF (RC /= 0) CALL FAIL('BQ_GET_FIELD', RC)
      CALL STORE_SNAPSHOTS(I1)
    END IF
    I0 = I1 + 1
  END DO

  CALL BQ_DESTROY(HB)

CONTAINS

  SUBROUTINE PUT(A, K)
    REAL(PP), INTENT(IN) :: A(:,:)
    INTEGER, INTENT(IN)  :: K
    PBUF(OFF:OFF+SIZE(A)-1) = RESHAPE(A, [SIZE(A)])
    COEF(K) = C_LOC(PBUF(OFF))
    OFF = OFF + SIZE(A)
  END SUBROUTINE PUT

  ! Receiver files of steps I0..I1 from the chunk's samples TR(:,J,S): per
  ! step the lines STORE_SEISMOGRAMS (store_seismograms.F90:54-100, 140-153)
  ! and STORE_SEISMOGRAMS_PRESSURE (store_seismograms_pressure.F90:45-81)
  ! append, with the derivatives formed from the previous samples SSU, SSW,
  ! SSQX, SSQZ as there; each file is opened once for the whole chunk.
  SUBROUTINE STORE_CHUNK(I0, I1)
    INTEGER, INTENT(IN) :: I0, I1
    INTEGER  :: J, T, S
    REAL(PP) :: DU, DW, DQX, DQZ, P, X
    LOGICAL  :: TLQ
    TLQ = KEY_TLV .AND. KEY_TLQ
    DO J = 1, MR
      IF (KEY_TLV) OPEN (21, FILE='REC_UW_'//TRIM(STR(J))//'.asc', STATUS='UNKNOWN',  &
                         POSITION='APPEND')
      IF (TLQ) OPEN (22, FILE='REC_Q_'//TRIM(STR(J))//'.asc', STATUS='UNKNOWN',       &
                     POSITION='APPEND')
      IF (KEY_TLP) OPEN (27, FILE='REC_STRESS_'//TRIM(STR(J))//'.asc', STATUS='UNKNOWN', &
                         POSITION='APPEND')
      DO T = I0, I1
        S = T - I0 + 1
        DU = 0.
        DW = 0.
        DQX = 0.
        DQZ = 0.
        IF (KEY_TLV .AND. INSIDE(IRECX(J), LRECX(J))) THEN
          DU = (TR(1, J, S) - SSU(J))/DT
          SSU(J) = TR(1, J, S)
          IF (TLQ) THEN
            DQX = (TR(3, J, S) - SSQX(J))/DT
            SSQX(J) = TR(3, J, S)
          END IF
        END IF
        IF (KEY_TLV .AND. INSIDE(IRECZ(J), LRECZ(J))) THEN
          DW = (TR(2, J, S) - SSW(J))/DT
          SSW(J) = TR(2, J, S)
          IF (TLQ) THEN
            DQZ = (TR(4, J, S) - SSQZ(J))/DT
            SSQZ(J) = TR(4, J, S)
          END IF
        END IF
        IF (KEY_TLV) WRITE (21, '(5(1X,E13.6))') DT*(T), SSU(J), SSW(J), DU, DW
        IF (TLQ) WRITE (22, '(5(1X,E13.6))') DT*(T), SSQX(J), SSQZ(J), DQX, DQZ
        IF (KEY_TLP) THEN
          P = 0.
          X = 0.
          IF (INSIDE(IRECZ(J), LRECX(J))) THEN
            P = TR(5, J, S)
            X = TR(6, J, S)
          END IF
          WRITE (27, '(3(1X,E13.6))') DT*(T), P, X
        END IF
      END DO
      IF (KEY_TLV) CLOSE (21)
      IF (TLQ) CLOSE (22)
      IF (KEY_TLP) CLOSE (27)
    END DO
  END SUBROUTINE STORE_CHUNK

  LOGICAL FUNCTION INSIDE(I, L)
    INTEGER, INTENT(IN) :: I, L
    INSIDE = X_MIN <= I .AND. I <= X_MAX .AND. Z_MIN <= L .AND. L <= Z_MAX
  END FUNCTION INSIDE

  ! Error path of the reference: message to the log unit, then STOP
  ! (tl_alloc.F90:39-42).
  SUBROUTINE FAIL(WHAT, CODE)
    CHARACTER(LEN=*), INTENT(IN) :: WHAT
    INTEGER, INTENT(IN) :: CODE
    WRITE (11,*) '  ', WHAT, ' FAILED, CODE= ', CODE, ' ', TRIM(BQ_MESSAGE(HB))
    WRITE (*,*)  '  ', WHAT, ' FAILED, CODE= ', CODE, ' ', TRIM(BQ_MESSAGE(HB))
    STOP
  END SUBROUTINE FAIL

END SUBROUTINE TIME_LOOP
