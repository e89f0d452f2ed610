C=======================================================================
C     Drop-in SUBROUTINE STEP (replaces 3dmhdsub.f:687-1176).
C
C     Same name, no arguments and the same COMMON blocks as the
C     reference; the step itself runs on the GPU through the C-ABI of
C     include/rast3dmhd.h (module RAST3DMHD_C).  Link this object ahead of
C     (or instead of) the reference's STEP; the driver is unchanged.
C
C     First call: describes the box (PARAMETER block and COMMON scalars),
C     uploads the metric arrays and the state, and seeds the library's
C     clock with NIT, TIMC, TIMI.
C     Every call: one RK3 step on the device.  DT and UMACH come back as
C     STEP leaves them in /CTIM/ and /TRACE/, and NIT, TIMC, TIMT advance
C     as in the reference (sub:1171-1173).  The prognostic arrays of /BIG/
C     are copied back whenever the driver reads them: dump steps
C     (MOD(NIT,NSTEP0).EQ.0), the last step (NIT.GE.NTOTAL) and the
C     DT.LT.1.0E-08 abort path (3dmhd.f:924-1009, 1053-1153).  Between
C     those, the device holds the authoritative state.
C
C     The whole domain is resident on one GPU: the parameter block must
C     have NPEY=NPEZ=1.  Errors print the library's message and STOP, as
C     the reference does for invalid settings.
C=======================================================================
      SUBROUTINE STEP
      USE ISO_C_BINDING
      USE RAST3DMHD_C
      INCLUDE '3dmhdparam.f'
C
      DIMENSION RU(NX,NY,NZ),RV(NX,NY,NZ),RW(NX,NY,NZ),RO(NX,NY,NZ)
     2         ,TT(NX,NY,NZ)
      DIMENSION UU(NX,NY,NZ),VV(NX,NY,NZ),WW(NX,NY,NZ)
      DIMENSION FU(NX,NY,NZ),FV(NX,NY,NZ),FW(NX,NY,NZ),FR(NX,NY,NZ)
     2         ,FT(NX,NY,NZ)
      DIMENSION ZRU(NX,NY,NZ),ZRV(NX,NY,NZ),ZRW(NX,NY,NZ)
     2         ,ZRO(NX,NY,NZ),ZTT(NX,NY,NZ)
      DIMENSION WW1(NX,NY,NZ),WW2(NX,NY,NZ),WW3(NX,NY,NZ)
      DIMENSION BX(NX,NY,NZ),BY(NX,NY,NZ),BZ(NX,NY,NZ)
      DIMENSION ZBX(NX,NY,NZ),ZBY(NX,NY,NZ),ZBZ(NX,NY,NZ)
      DIMENSION EXX(NX),DXXDX(NX),D2XXDX2(NX),DDX(NX)
      DIMENSION WYY(NY),DYYDY(NY),D2YYDY2(NY),DDY(NY)
      DIMENSION ZEE(NZ),DZZDZ(NZ),D2ZZDZ2(NZ),DDZ(NZ)
      DIMENSION RKAPA(NZ),DKAPA(NZ)
      DIMENSION SP1(IPAD),SP2(IPAD),SP3(IPAD),SP4(IPAD),SP5(IPAD)
     2         ,SP6(IPAD),SP7(IPAD),SP8(IPAD),SP9(IPAD),SP10(IPAD)
     3         ,SP11(IPAD),SP12(IPAD),SP13(IPAD),SP14(IPAD),SP15(IPAD)
     4         ,SP16(IPAD),SP17(IPAD),SP18(IPAD),SP19(IPAD),SP20(IPAD)
     5         ,SP21(IPAD),SP22(IPAD),SP23(IPAD),SP24(IPAD),SP25(IPAD)
     6         ,SP26(IPAD)
C
      COMMON/BIG/RU,SP1,RV,SP2,RW,SP3,RO,SP4,TT,SP5,UU,SP6,VV,SP7,WW
     2          ,SP8,FU,SP9,FV,SP10,FW,SP11,FR,SP12,FT,SP13
     3          ,ZRU,SP14,ZRV,SP15,ZRW,SP16,ZRO,SP17,ZTT
     4          ,SP18,WW1,SP19,WW2,SP20,WW3
     5          ,SP21,BX,SP22,BY,SP23,BZ,SP24,ZBX,SP25,ZBY,SP26,ZBZ
      COMMON/AJACOBI/EXX,DXXDX,D2XXDX2,DDX,WYY,DYYDY,D2YYDY2,DDY
     2              ,ZEE,DZZDZ,D2ZZDZ2,DDZ
      COMMON/BCT/IXC,IYC,IZC,ITC,IBC
      COMMON/ITER/NTOTAL,NSTEP0,NIT
      COMMON/CPAR/CV,OCV,ORE,RE,REPR,THETA,GRAV,AMPT,SF,GAMMA
      COMMON/CROT/OMX,OMZ
      COMMON/CMAG/ORM,RM,OBETA,AMPB,BFH,BZP
      COMMON/CPEN/PZP,SIGMA,RKAPST,TB,RKAPA,DKAPA,RKAPM
      COMMON/CPER/TP,XP,YP,ZP,TC,QFH,HH
      COMMON/GRID/DD,HX,H2X,HY,H2Y,HZ,H2Z,C13,C23,C43
      COMMON/CTIM/DT,TIMT,TIMC,TIMI
      COMMON/TRACE/UMACH
      COMMON/RUNGKU/GAM1,GAM2,GAM3,ZETA1,ZETA2
C
      TYPE(C_PTR), SAVE :: H
      LOGICAL, SAVE :: READY = .FALSE.
      TYPE(MHD_DESC) :: D
      INTEGER(C_INT) :: IRC
      REAL(C_DOUBLE) :: DTC, UMC
      CHARACTER(LEN=16) :: PREC
      INTEGER :: PLEN, PST, IB, JB, KB
      REAL(C_DOUBLE) :: RO0(NZ), TT0(NZ)
C
      IF (.NOT.READY) THEN
         IF (NPEY*NPEZ.NE.1) THEN
            WRITE(6,*)'STEP: the GPU step needs NPEY=NPEZ=1'
            STOP
         ENDIF
         IF ((IXC.NE.0).OR.(IYC.NE.0)) THEN
            WRITE(6,*)'FLUXES: Non-periodic horizontal boundaries'
            STOP
         ENDIF
         D%NPX=NPX
         D%NPY=NPY
         D%NPZ=NPZ
         D%LMAG=0
         IF (LMAG) D%LMAG=1
         D%LROT=0
         IF (LROT) D%LROT=1
         D%IZCON=IZC
         D%ITCON=ITC
         D%IBCON=IBC
C        RAST3DMHD_PRECISION=mixed selects the mixed-precision form
         D%PRECISION=MHD_FP64
         CALL GET_ENVIRONMENT_VARIABLE('RAST3DMHD_PRECISION',PREC,
     2                                 PLEN,PST)
         IF ((PST.EQ.0).AND.(PLEN.GT.0)) THEN
            IF (PREC(1:PLEN).EQ.'mixed') THEN
               D%PRECISION=MHD_MIXED
            ELSE IF (PREC(1:PLEN).NE.'fp64') THEN
               WRITE(6,*)'STEP: RAST3DMHD_PRECISION is fp64 or mixed'
               STOP
            ENDIF
         ENDIF
         D%HX=HX
         D%H2X=H2X
         D%HY=HY
         D%H2Y=H2Y
         D%HZ=HZ
         D%H2Z=H2Z
         D%C13=C13
         D%C23=C23
         D%C43=C43
         D%GAM1=GAM1
         D%GAM2=GAM2
         D%GAM3=GAM3
         D%ZETA1=ZETA1
         D%ZETA2=ZETA2
         D%CV=CV
         D%OCV=OCV
         D%ORE=ORE
         D%RE=RE
         D%REPR=REPR
         D%THETA=THETA
         D%GRAV=GRAV
         D%SF=SF
         D%GAMMA=GAMMA
         D%OMX=OMX
         D%OMZ=OMZ
         D%ORM=ORM
         D%RM=RM
         D%OBETA=OBETA
         D%TB=TB
         D%TP=TP
         D%DZF=1.0E00/FLOAT(NPZ-1)
         IRC=MHD_CREATE(D,H)
         IF (IRC.NE.0) CALL MHD_FAIL(C_NULL_PTR,'STEP: mhd_create',IRC)
         IRC=MHD_SET_METRICS(H,DXXDX,D2XXDX2,DDX,DYYDY,D2YYDY2,DDY,
     2                       DZZDZ,D2ZZDZ2,DDZ,RKAPA,DKAPA)
         IF (IRC.NE.0) CALL MHD_FAIL(H,'STEP: mhd_set_metrics',IRC)
         IF (D%PRECISION.EQ.MHD_MIXED) THEN
C           background of the mixed form: horizontal means of rho and T
            DO KB=1,NZ
               RO0(KB)=0.0D0
               TT0(KB)=0.0D0
               DO JB=2,NY-1
                  DO IB=2,NX-1
                     RO0(KB)=RO0(KB)+RO(IB,JB,KB)
                     TT0(KB)=TT0(KB)+TT(IB,JB,KB)
                  ENDDO
               ENDDO
               RO0(KB)=RO0(KB)/DBLE((NX-2)*(NY-2))
               TT0(KB)=TT0(KB)/DBLE((NX-2)*(NY-2))
            ENDDO
            IRC=MHD_SET_BACKGROUND(H,RO0,TT0)
            IF (IRC.NE.0) CALL MHD_FAIL(H,'STEP: mhd_set_background',
     2                                  IRC)
         ENDIF
         IRC=MHD_UPLOAD(H,RU,RV,RW,RO,TT,BX,BY,BZ)
         IF (IRC.NE.0) CALL MHD_FAIL(H,'STEP: mhd_upload',IRC)
         IRC=MHD_SET_CLOCK(H,INT(NIT,C_INT),TIMC,TIMI)
         IF (IRC.NE.0) CALL MHD_FAIL(H,'STEP: mhd_set_clock',IRC)
         READY=.TRUE.
      ENDIF
C
      IRC=MHD_STEP(H,DTC,UMC)
      IF (IRC.NE.0) CALL MHD_FAIL(H,'STEP: mhd_step',IRC)
      DT=DTC
      UMACH=UMC
      NIT=NIT+1
      TIMC=TIMC+DT
      TIMT=TIMI+TIMC
C
      IF ((MOD(NIT,NSTEP0).EQ.0).OR.(NIT.GE.NTOTAL)
     2    .OR.(DT.LT.1.0E-08)) THEN
         IRC=MHD_DOWNLOAD(H,RU,RV,RW,RO,TT,BX,BY,BZ)
         IF (IRC.NE.0) CALL MHD_FAIL(H,'STEP: mhd_download',IRC)
      ENDIF
      RETURN
      END
