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
C     Decomposition.  The driver runs on NPE=NPEY*NPEZ MPI ranks, as the
C     reference requires (3dmhdparam.f:11; rank MYPE = MYPEY+NPEY*MYPEZ
C     holds rows MYPEY*NRY+1.. and planes MYPEZ*NRZ+1.., 3dmhd.f:87-88).
C     The whole domain is resident on the GPU of rank 0: on the first
C     call every rank's slab of /BIG/ and of the y and z metrics is
C     gathered there, and rank 0 alone creates the handle and steps.  DT
C     and UMACH are broadcast every step.  Where the driver reads /BIG/,
C     rank 0 downloads the whole domain and scatters every rank's slab
C     back with its ghost layers as COMMUNICATE leaves them (neighbour
C     rows and planes, periodic wrap; the untouched physical z ghosts).
C     With NPE=1 no MPI call is made and /BIG/ is used directly.
C
C     The optional branches of FLUXES/BCON (ID, LREM, LSHR, RLAX, ITCON
C     2..4, TP) are passed on with the coordinates EXX, WYY, ZEE; the
C     library refuses those it does not have in the chosen form.  Errors
C     print the library's message and stop, as the reference does for
C     invalid settings.
C=======================================================================
      SUBROUTINE STEP
      USE ISO_C_BINDING
      USE RAST3DMHD_C
      INCLUDE '3dmhdparam.f'
      INCLUDE 'mpif.h'
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
      COMMON/COMMUN/MYPE,MYPEY,MYPEZ,MPISIZE
      COMMON/RELAX/TSTART,TOFF,RLAX
      COMMON/SPECIALBOUND/TU,DZTB,DZTU
      COMMON/BOUNDS/XMAX,YMAX,ZMAX
C
      TYPE(C_PTR), SAVE :: H
      LOGICAL, SAVE :: READY = .FALSE.
C     rank 0 of a decomposed run: the whole padded domain, /BIG/ order
C     RU,RV,RW,RO,TT,BX,BY,BZ, and the y and z metric arrays
      REAL(C_DOUBLE), ALLOCATABLE, SAVE :: G(:,:,:,:)
      REAL(C_DOUBLE), ALLOCATABLE, SAVE :: GY(:,:), GZ(:,:)
C     (GY: DYYDY, D2YYDY2, DDY, WYY; GZ: DZZDZ, D2ZZDZ2, DDZ, RKAPA, DKAPA,
C     ZEE)
      REAL(C_DOUBLE), ALLOCATABLE :: BUF(:,:,:,:), MV(:), MB(:,:)
      TYPE(MHD_DESC) :: D
      INTEGER(C_INT) :: IRC
      REAL(C_DOUBLE) :: DTC, UMC, V3(3), TBG(1)
      CHARACTER(LEN=16) :: PREC
      CHARACTER(LEN=40) :: WHAT
      INTEGER :: PLEN, PST, NF, IERR, IR, PY, PZ, J, K, JG, KG
      INTEGER :: NRB, IOS
      INTEGER :: NSEND, NMV
      LOGICAL :: DUMP, FL
C
      NSEND=NX*NY*NZ
      IF (.NOT.READY) THEN
         IF ((IXC.NE.0).OR.(IYC.NE.0)) THEN
            WRITE(6,*)'FLUXES: Non-periodic horizontal boundaries'
            CALL SHIM_STOP
         ENDIF
         NF=5
         IF (LMAG) NF=8
         TBG(1)=TB
         IF (NPE.GT.1) THEN
C           TB is the bottom ranks' value on a restart (3dmhd.f:354-356)
            CALL MPI_BCAST(TBG,1,MPI_DOUBLE_PRECISION,NPE-1,
     2                     MPI_COMM_WORLD,IERR)
            IF (MYPE.EQ.0) THEN
               ALLOCATE(G(NX,NPY+2,NPZ+2,NF),BUF(NX,NY,NZ,0:NPE-1))
               ALLOCATE(GY(NPY+2,4),GZ(NPZ+2,6))
            ELSE
               ALLOCATE(BUF(1,1,1,0:0))
            ENDIF
            CALL GATHER3(RU,1)
            CALL GATHER3(RV,2)
            CALL GATHER3(RW,3)
            CALL GATHER3(RO,4)
            CALL GATHER3(TT,5)
            IF (LMAG) THEN
               CALL GATHER3(BX,6)
               CALL GATHER3(BY,7)
               CALL GATHER3(BZ,8)
            ENDIF
            CALL GATHER_METRICS
         ENDIF
         IRC=0
         IF (MYPE.EQ.0) CALL SHIM_CREATE
         V3(3)=DBLE(IRC)
         IF (NPE.GT.1) CALL MPI_BCAST(V3(3:3),1,MPI_DOUBLE_PRECISION,
     2                                0,MPI_COMM_WORLD,IERR)
         IF (V3(3).NE.0.0D0) THEN
            IF (MYPE.EQ.0) CALL MHD_REPORT(H,TRIM(WHAT),
     2                                     INT(V3(3),C_INT))
            CALL SHIM_STOP
         ENDIF
         READY=.TRUE.
      ENDIF
C
      IF (MYPE.EQ.0) THEN
         IRC=MHD_STEP(H,DTC,UMC)
         V3(1)=DTC
         V3(2)=UMC
         V3(3)=DBLE(IRC)
      ENDIF
      IF (NPE.GT.1) CALL MPI_BCAST(V3,3,MPI_DOUBLE_PRECISION,0,
     2                             MPI_COMM_WORLD,IERR)
      IF (V3(3).NE.0.0D0) THEN
         IF (MYPE.EQ.0) CALL MHD_REPORT(H,'STEP: mhd_step',
     2                                  INT(V3(3),C_INT))
         CALL SHIM_STOP
      ENDIF
      DT=V3(1)
      UMACH=V3(2)
      NIT=NIT+1
      TIMC=TIMC+DT
      TIMT=TIMI+TIMC
C
      DUMP=(MOD(NIT,NSTEP0).EQ.0).OR.(NIT.GE.NTOTAL)
     2     .OR.(DT.LT.1.0E-08)
      IF (DUMP) THEN
         IF (NPE.EQ.1) THEN
            IRC=MHD_DOWNLOAD(H,RU,RV,RW,RO,TT,BX,BY,BZ)
            IF (IRC.NE.0) CALL MHD_FAIL(H,'STEP: mhd_download',IRC)
         ELSE
            IF (MYPE.EQ.0) THEN
               IF (LMAG) THEN
                  IRC=MHD_DOWNLOAD(H,G(:,:,:,1),G(:,:,:,2),G(:,:,:,3),
     2                 G(:,:,:,4),G(:,:,:,5),G(:,:,:,6),G(:,:,:,7),
     3                 G(:,:,:,8))
               ELSE
C                 (BX,BY,BZ of /BIG/ are not touched without LMAG)
                  IRC=MHD_DOWNLOAD(H,G(:,:,:,1),G(:,:,:,2),G(:,:,:,3),
     2                 G(:,:,:,4),G(:,:,:,5),BX,BY,BZ)
               ENDIF
               V3(3)=DBLE(IRC)
            ENDIF
            CALL MPI_BCAST(V3(3:3),1,MPI_DOUBLE_PRECISION,0,
     2                     MPI_COMM_WORLD,IERR)
            IF (V3(3).NE.0.0D0) THEN
               IF (MYPE.EQ.0) CALL MHD_REPORT(H,'STEP: mhd_download',
     2                                        INT(V3(3),C_INT))
               CALL SHIM_STOP
            ENDIF
            ALLOCATE(BUF(NX,NY,NZ,0:MERGE(NPE-1,0,MYPE.EQ.0)))
            CALL SCATTER3(RU,1)
            CALL SCATTER3(RV,2)
            CALL SCATTER3(RW,3)
            CALL SCATTER3(RO,4)
            CALL SCATTER3(TT,5)
            IF (LMAG) THEN
               CALL SCATTER3(BX,6)
               CALL SCATTER3(BY,7)
               CALL SCATTER3(BZ,8)
            ENDIF
            DEALLOCATE(BUF)
         ENDIF
      ENDIF
      RETURN
C
      CONTAINS
C
C     The reference's stop on invalid settings: print, finalise MPI
C     (every rank reaches this together), stop.
      SUBROUTINE SHIM_STOP
      INTEGER :: IE
      CALL MPI_FINALIZED(FL,IE)
      IF (.NOT.FL) THEN
         CALL MPI_INITIALIZED(FL,IE)
         IF (FL) CALL MPI_FINALIZE(IE)
      ENDIF
      STOP
      END SUBROUTINE SHIM_STOP
C
C     Rank 0: local array LOC of every rank into G(:,:,:,F); interior
C     rows and planes from their owner, the outer ghost rows and planes
C     from the edge ranks.
      SUBROUTINE GATHER3(LOC,F)
      REAL*8 :: LOC(NX,NY,NZ)
      INTEGER :: F
      CALL MPI_GATHER(LOC,NSEND,MPI_DOUBLE_PRECISION,BUF,NSEND,
     2                MPI_DOUBLE_PRECISION,0,MPI_COMM_WORLD,IERR)
      IF (MYPE.NE.0) RETURN
      DO IR=0,NPE-1
         PY=MOD(IR,NPEY)
         PZ=IR/NPEY
         DO K=1,NZ
            IF ((K.EQ.1).AND.(PZ.GT.0)) CYCLE
            IF ((K.EQ.NZ).AND.(PZ.LT.NPEZ-1)) CYCLE
            KG=PZ*NRZ+K
            DO J=1,NY
               IF ((J.EQ.1).AND.(PY.GT.0)) CYCLE
               IF ((J.EQ.NY).AND.(PY.LT.NPEY-1)) CYCLE
               JG=PY*NRY+J
               G(:,JG,KG,F)=BUF(:,J,K,IR)
            ENDDO
         ENDDO
      ENDDO
      END SUBROUTINE GATHER3
C
C     Rank 0 to every rank: its window of G(:,:,:,F), ghosts included.
      SUBROUTINE SCATTER3(LOC,F)
      REAL*8 :: LOC(NX,NY,NZ)
      INTEGER :: F
      IF (MYPE.EQ.0) THEN
         DO IR=0,NPE-1
            PY=MOD(IR,NPEY)
            PZ=IR/NPEY
            BUF(:,:,:,IR)=G(:,PY*NRY+1:PY*NRY+NY,PZ*NRZ+1:PZ*NRZ+NZ,F)
         ENDDO
      ENDIF
      CALL MPI_SCATTER(BUF,NSEND,MPI_DOUBLE_PRECISION,LOC,NSEND,
     2                 MPI_DOUBLE_PRECISION,0,MPI_COMM_WORLD,IERR)
      END SUBROUTINE SCATTER3
C
C     Rank 0: the y metrics (DYYDY, D2YYDY2, DDY, WYY) and z metrics
C     (DZZDZ, D2ZZDZ2, DDZ, RKAPA, DKAPA, ZEE) of the whole domain from the
C     slabs.
      SUBROUTINE GATHER_METRICS
      NMV=4*NY+6*NZ
      ALLOCATE(MV(NMV))
      IF (MYPE.EQ.0) THEN
         ALLOCATE(MB(NMV,0:NPE-1))
      ELSE
         ALLOCATE(MB(1,0:0))
      ENDIF
      MV(1:NY)=DYYDY
      MV(NY+1:2*NY)=D2YYDY2
      MV(2*NY+1:3*NY)=DDY
      MV(3*NY+1:4*NY)=WYY
      MV(4*NY+1:4*NY+NZ)=DZZDZ
      MV(4*NY+NZ+1:4*NY+2*NZ)=D2ZZDZ2
      MV(4*NY+2*NZ+1:4*NY+3*NZ)=DDZ
      MV(4*NY+3*NZ+1:4*NY+4*NZ)=RKAPA
      MV(4*NY+4*NZ+1:4*NY+5*NZ)=DKAPA
      MV(4*NY+5*NZ+1:4*NY+6*NZ)=ZEE
      CALL MPI_GATHER(MV,NMV,MPI_DOUBLE_PRECISION,MB,NMV,
     2                MPI_DOUBLE_PRECISION,0,MPI_COMM_WORLD,IERR)
      IF (MYPE.EQ.0) THEN
         DO IR=0,NPE-1
            PY=MOD(IR,NPEY)
            PZ=IR/NPEY
            DO J=1,NY
               IF ((J.EQ.1).AND.(PY.GT.0)) CYCLE
               IF ((J.EQ.NY).AND.(PY.LT.NPEY-1)) CYCLE
               JG=PY*NRY+J
               GY(JG,1)=MB(J,IR)
               GY(JG,2)=MB(NY+J,IR)
               GY(JG,3)=MB(2*NY+J,IR)
               GY(JG,4)=MB(3*NY+J,IR)
            ENDDO
            DO K=1,NZ
               IF ((K.EQ.1).AND.(PZ.GT.0)) CYCLE
               IF ((K.EQ.NZ).AND.(PZ.LT.NPEZ-1)) CYCLE
               KG=PZ*NRZ+K
               GZ(KG,1)=MB(4*NY+K,IR)
               GZ(KG,2)=MB(4*NY+NZ+K,IR)
               GZ(KG,3)=MB(4*NY+2*NZ+K,IR)
               GZ(KG,4)=MB(4*NY+3*NZ+K,IR)
               GZ(KG,5)=MB(4*NY+4*NZ+K,IR)
               GZ(KG,6)=MB(4*NY+5*NZ+K,IR)
            ENDDO
         ENDDO
      ENDIF
      DEALLOCATE(MV,MB)
      END SUBROUTINE GATHER_METRICS
C
C     Rank 0: handle, metrics, background (mixed form), state, clock.
C     A failure leaves its code in IRC and the failing call in WHAT.
      SUBROUTINE SHIM_CREATE
      REAL(C_DOUBLE) :: RO0(NPZ+2), TT0(NPZ+2)
      INTEGER :: IB, JB, KB
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
C     RAST3DMHD_PRECISION=mixed selects the mixed-precision form
      D%PRECISION=MHD_FP64
      CALL GET_ENVIRONMENT_VARIABLE('RAST3DMHD_PRECISION',PREC,
     2                              PLEN,PST)
      IF ((PST.EQ.0).AND.(PLEN.GT.0)) THEN
         IF (PREC(1:PLEN).EQ.'mixed') THEN
            D%PRECISION=MHD_MIXED
         ELSE IF (PREC(1:PLEN).NE.'fp64') THEN
            WRITE(6,*)'STEP: RAST3DMHD_PRECISION is fp64 or mixed'
            IRC=1
            WHAT='STEP: RAST3DMHD_PRECISION'
            RETURN
         ENDIF
      ENDIF
C     RAST3DMHD_REBASE=N re-centres the mixed form's background on the
C     horizontal means every N steps (mhd_set_rebase_interval)
      NRB=0
      CALL GET_ENVIRONMENT_VARIABLE('RAST3DMHD_REBASE',PREC,PLEN,PST)
      IF ((PST.EQ.0).AND.(PLEN.GT.0)) THEN
         READ(PREC(1:PLEN),*,IOSTAT=IOS) NRB
         IF ((IOS.NE.0).OR.(NRB.LT.0)) THEN
            WRITE(6,*)'STEP: RAST3DMHD_REBASE is a number of steps'
            IRC=1
            WHAT='STEP: RAST3DMHD_REBASE'
            RETURN
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
      D%TB=TBG(1)
      D%TP=TP
      D%DZF=1.0E00/FLOAT(NPZ-1)
C     optional branches (3dmhdparam.f, /SPECIALBOUND/, /RELAX/, /CPER/,
C     /BOUNDS/) and the decomposition whose summation order they follow
      D%ID=ID
      D%LREM=0
      IF (LREM) D%LREM=1
      D%LSHR=0
      IF (LSHR) D%LSHR=1
      D%NPEY=NPEY
      D%NPEZ=NPEZ
      D%DH=DH
      D%DV=DV
      D%TU=TU
      D%DZTU=DZTU
      D%DZTB=DZTB
      D%RLAX=RLAX
      D%TSTART=TSTART
      D%TOFF=TOFF
      D%XP=XP
      D%YP=YP
      D%ZP=ZP
      D%TC=TC
      D%QFH=QFH
      D%HH=HH
      D%ZMAX=ZMAX
      WHAT='STEP: mhd_create'
      IRC=MHD_CREATE(D,H)
      IF (IRC.NE.0) GOTO 9
      WHAT='STEP: mhd_set_metrics'
      IF (NPE.EQ.1) THEN
         IRC=MHD_SET_METRICS(H,DXXDX,D2XXDX2,DDX,DYYDY,D2YYDY2,DDY,
     2                       DZZDZ,D2ZZDZ2,DDZ,RKAPA,DKAPA)
      ELSE
         IRC=MHD_SET_METRICS(H,DXXDX,D2XXDX2,DDX,GY(:,1),GY(:,2),
     2                       GY(:,3),GZ(:,1),GZ(:,2),GZ(:,3),GZ(:,4),
     3                       GZ(:,5))
      ENDIF
      IF (IRC.NE.0) GOTO 9
      WHAT='STEP: mhd_set_coordinates'
      IF (NPE.EQ.1) THEN
         IRC=MHD_SET_COORDINATES(H,EXX,WYY,ZEE)
      ELSE
         IRC=MHD_SET_COORDINATES(H,EXX,GY(:,4),GZ(:,6))
      ENDIF
      IF (IRC.NE.0) GOTO 9
      IF (D%PRECISION.EQ.MHD_MIXED) THEN
C        background of the mixed form: horizontal means of rho and T
         DO KB=1,NPZ+2
            RO0(KB)=0.0D0
            TT0(KB)=0.0D0
            DO JB=2,NPY+1
               DO IB=2,NX-1
                  IF (NPE.EQ.1) THEN
                     RO0(KB)=RO0(KB)+RO(IB,JB,KB)
                     TT0(KB)=TT0(KB)+TT(IB,JB,KB)
                  ELSE
                     RO0(KB)=RO0(KB)+G(IB,JB,KB,4)
                     TT0(KB)=TT0(KB)+G(IB,JB,KB,5)
                  ENDIF
               ENDDO
            ENDDO
            RO0(KB)=RO0(KB)/DBLE(NPX*NPY)
            TT0(KB)=TT0(KB)/DBLE(NPX*NPY)
         ENDDO
         WHAT='STEP: mhd_set_background'
         IRC=MHD_SET_BACKGROUND(H,RO0,TT0)
         IF (IRC.NE.0) GOTO 9
         IF (NRB.GT.0) THEN
            WHAT='STEP: mhd_set_rebase_interval'
            IRC=MHD_SET_REBASE_INTERVAL(H,NRB)
            IF (IRC.NE.0) GOTO 9
         ENDIF
      ENDIF
      WHAT='STEP: mhd_upload'
      IF (NPE.EQ.1) THEN
         IRC=MHD_UPLOAD(H,RU,RV,RW,RO,TT,BX,BY,BZ)
      ELSE IF (LMAG) THEN
         IRC=MHD_UPLOAD(H,G(:,:,:,1),G(:,:,:,2),G(:,:,:,3),G(:,:,:,4),
     2                  G(:,:,:,5),G(:,:,:,6),G(:,:,:,7),G(:,:,:,8))
      ELSE
         IRC=MHD_UPLOAD(H,G(:,:,:,1),G(:,:,:,2),G(:,:,:,3),G(:,:,:,4),
     2                  G(:,:,:,5),BX,BY,BZ)
      ENDIF
      IF (IRC.NE.0) GOTO 9
      WHAT='STEP: mhd_set_clock'
      IRC=MHD_SET_CLOCK(H,INT(NIT,C_INT),TIMC,TIMI)
    9 CONTINUE
      END SUBROUTINE SHIM_CREATE
      END
