C=======================================================================
C     Golden-vector harness (TEST INFRASTRUCTURE ONLY).
C
C     Drives the reference's own set-up and time-step routines -- XJACOBI,
C     YJACOBI, ZJACOBI, STATIC, COMMUNICATE, FLUXES and STEP, linked from
C     the reference's 3dmhdsub.o -- on a small box under MPI, and dumps the
C     per-rank state so that tests can pin the C oracle and the GPU path
C     against the reference itself.  The COMMON set-up below mirrors the
C     reference main program (3dmhd.f:105-279, 378-379).
C
C     Input file harness.in (read by every rank):
C       line 1: NSTEPS
C       line 2: NDUMP, then NDUMP step numbers after which to dump
C       line 3: AMPU AMPBP ISEED  (velocity and field perturbation)
C     The perturbation is a deterministic integer hash of the GLOBAL
C     indices, so the initial state is independent of the decomposition.
C=======================================================================
      PROGRAM REFHARNESS
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
      DIMENSION IDUMP(64)
      REAL*8, ALLOCATABLE :: SAV(:,:,:,:)
      CHARACTER*64 FNAME
C
      COMMON/BIG/RU,RV,RW,RO,TT,UU,VV,WW,FU,FV,FW,FR,FT
     2          ,ZRU,ZRV,ZRW,ZRO,ZTT,WW1,WW2,WW3
     3          ,BX,BY,BZ,ZBX,ZBY,ZBZ
      COMMON/AJACOBI/EXX,DXXDX,D2XXDX2,DDX,WYY,DYYDY,D2YYDY2,DDY
     2              ,ZEE,DZZDZ,D2ZZDZ2,DDZ
      COMMON/BCT/IXC,IYC,IZC,ITC,IBC
      COMMON/ITER/NTOTAL,NSTEP0,NIT
      COMMON/CPAR/CV,OCV,ORE,RE,REPR,THETA,GRAV,AMPT,SF,GAMMA
      COMMON/CROT/OMX,OMZ
      COMMON/CMAG/ORM,RM,OBETA,AMPB,BFH,BZP
      COMMON/CPEN/PZP,SIGMA,RKAPST,TB,RKAPA,DKAPA,RKAPM
      COMMON/CPER/TP,XP,YP,ZP,TC,QFH,HH
      COMMON/BOUNDS/XMAX,YMAX,ZMAX
      COMMON/GRID/DD,HX,H2X,HY,H2Y,HZ,H2Z,C13,C23,C43
      COMMON/CTIM/DT,TIMT,TIMC,TIMI
      COMMON/TRACE/UMACH
      COMMON/RUNGKU/GAM1,GAM2,GAM3,ZETA1,ZETA2
      COMMON/COMMUN/MYPE,MYPEY,MYPEZ,MPISIZE
      COMMON/RELAX/TSTART,TOFF,RLAX
      COMMON/SPECIALBOUND/TU,DZTB,DZTU
C
      CALL MPI_INIT(IERR)
      CALL MPI_COMM_RANK(MPI_COMM_WORLD,MYPE,IERR)
      CALL MPI_COMM_SIZE(MPI_COMM_WORLD,NN,IERR)
      IF (NN.NE.NPE) THEN
         WRITE(6,*)'HARNESS: rank count differs from NPE'
         CALL MPI_ABORT(MPI_COMM_WORLD,1,IERR)
      ENDIF
      MYPEY=MOD(MYPE,NPEY)
      MYPEZ=MYPE/NPEY
      MPISIZE=MPI_DOUBLE_PRECISION
C
      OPEN(30,FILE='harness.in',STATUS='OLD')
      READ(30,*) NSTEPS
      READ(30,*) NDUMP,(IDUMP(N),N=1,NDUMP)
      READ(30,*) AMPU,AMPBP,ISEED
      CLOSE(30)
C
C     COMMON set-up of the reference main program.  The statements are
C     the reference's own (3dmhd.f:105-240: constants, SETUP, IPAR/PAR to
C     COMMON; 3dmhd.f:269-279: RKAPA, DKAPA), extracted from the reference
C     source at build time by build_ref.py into the two include files.
      INCLUDE 'drvset1.inc'
      TIMC=0.0E00
      CALL XJACOBI
      CALL YJACOBI
      CALL ZJACOBI
      INCLUDE 'drvset2.inc'
C     A first STATIC call without T noise or B layer yields the bare
C     polytrope profiles T(z), R(z) (one column), which let tests rebuild
C     the perturbed initial state exactly; the second call is the real one.
      SAVT=AMPT
      SAVB=AMPB
      AMPT=0.0D0
      AMPB=0.0D0
      CALL STATIC
      WRITE(FNAME,'(A,I4.4,A)') 'prof_r',MYPE,'.bin'
      OPEN(21,FILE=FNAME,ACCESS='STREAM',FORM='UNFORMATTED',
     2     STATUS='REPLACE')
      WRITE(21) TT(2,2,:),RO(2,2,:)
      CLOSE(21)
      AMPT=SAVT
      AMPB=SAVB
      RU=0.0D0
      RV=0.0D0
      RW=0.0D0
      RO=0.0D0
      TT=0.0D0
      BX=0.0D0
      BY=0.0D0
      BZ=0.0D0
      CALL STATIC
      TIMI=0.0E00
      WRITE(FNAME,'(A,I4.4,A)') 'bcol_r',MYPE,'.bin'
      OPEN(21,FILE=FNAME,ACCESS='STREAM',FORM='UNFORMATTED',
     2     STATUS='REPLACE')
      WRITE(21) BX(2,2,:)
      CLOSE(21)
C
C     Deterministic perturbation in global indices.  RW and BZ stay zero
C     on the boundary planes so the initial state satisfies the BCs there.
      DO 40 K=2,NZ-1
         KG=K+MYPEZ*NRZ
         DO 40 J=2,NY-1
            JG=J+MYPEY*NRY
            DO 40 I=2,NX-1
               RU(I,J,K)=RU(I,J,K)
     2                  +RO(I,J,K)*AMPU*HASHR(I,JG,KG,1,ISEED)
               RV(I,J,K)=RV(I,J,K)
     2                  +RO(I,J,K)*AMPU*HASHR(I,JG,KG,2,ISEED)
               IF ((KG.NE.2).AND.(KG.NE.NPZ+1)) THEN
                  RW(I,J,K)=RW(I,J,K)
     2                     +RO(I,J,K)*AMPU*HASHR(I,JG,KG,3,ISEED)
               ENDIF
               IF (LMAG) THEN
                  BX(I,J,K)=BX(I,J,K)+AMPBP*HASHR(I,JG,KG,4,ISEED)
                  BY(I,J,K)=BY(I,J,K)+AMPBP*HASHR(I,JG,KG,5,ISEED)
                  IF ((KG.NE.2).AND.(KG.NE.NPZ+1)) THEN
                     BZ(I,J,K)=BZ(I,J,K)+AMPBP*HASHR(I,JG,KG,6,ISEED)
                  ENDIF
               ENDIF
   40 CONTINUE
      CALL COMMUNICATE
      CV=1.0E00/(GAMMA-1.0E00)
      OCV=1.0E00/CV
C
C     Metadata: sizes, decomposition, every COMMON scalar STEP reads, and
C     the local metric arrays.
      WRITE(FNAME,'(A,I4.4,A)') 'meta_r',MYPE,'.bin'
      OPEN(21,FILE=FNAME,ACCESS='STREAM',FORM='UNFORMATTED',
     2     STATUS='REPLACE')
      WRITE(21) DBLE(NX),DBLE(NY),DBLE(NZ),DBLE(MYPEY),DBLE(MYPEZ),
     2   DBLE(NPEY),DBLE(NPEZ),DBLE(IZC),DBLE(ITC),DBLE(IBC)
      DZF=1.0E00/FLOAT(NPZ-1)
      WRITE(21) HX,H2X,HY,H2Y,HZ,H2Z,C13,C23,C43,
     2   GAM1,GAM2,GAM3,ZETA1,ZETA2,
     3   CV,OCV,ORE,RE,REPR,THETA,GRAV,SF,GAMMA,
     4   OMX,OMZ,ORM,RM,OBETA,TB,TP,DZF,AMPT
      WRITE(21) EXX,DXXDX,D2XXDX2,DDX,WYY,DYYDY,D2YYDY2,DDY,
     2   ZEE,DZZDZ,D2ZZDZ2,DDZ,RKAPA,DKAPA
      CLOSE(21)
C     Parameters of the optional FLUXES/BCON branches (density diffusion,
C     Rempel conduction, shear, relaxation, plume and heat loss).
      WRITE(FNAME,'(A,I4.4,A)') 'opt_r',MYPE,'.bin'
      OPEN(21,FILE=FNAME,ACCESS='STREAM',FORM='UNFORMATTED',
     2     STATUS='REPLACE')
      WRITE(21) DBLE(ID),DH,DV,TU,DZTU,DZTB,RLAX,TSTART,TOFF,
     2   XP,YP,ZP,TC,QFH,HH,XMAX,YMAX,ZMAX
      LOPT=0
      IF (LREM) LOPT=1
      WRITE(21) DBLE(LOPT)
      LOPT=0
      IF (LSHR) LOPT=1
      WRITE(21) DBLE(LOPT)
      CLOSE(21)
C
      WRITE(FNAME,'(A,I4.4,A)') 'init_r',MYPE,'.bin'
      OPEN(21,FILE=FNAME,ACCESS='STREAM',FORM='UNFORMATTED',
     2     STATUS='REPLACE')
      WRITE(21) RU,RV,RW,RO,TT,BX,BY,BZ
      CLOSE(21)
C
C     One FLUXES call on the initial state; the state is restored after.
      ALLOCATE(SAV(NX,NY,NZ,8))
      SAV(:,:,:,1)=RU
      SAV(:,:,:,2)=RV
      SAV(:,:,:,3)=RW
      SAV(:,:,:,4)=RO
      SAV(:,:,:,5)=TT
      SAV(:,:,:,6)=BX
      SAV(:,:,:,7)=BY
      SAV(:,:,:,8)=BZ
      CALL FLUXES
      WRITE(FNAME,'(A,I4.4,A)') 'flux_r',MYPE,'.bin'
      OPEN(21,FILE=FNAME,ACCESS='STREAM',FORM='UNFORMATTED',
     2     STATUS='REPLACE')
      WRITE(21) FU,FV,FW,FR,FT,WW1,WW2,WW3,UU,VV,WW,RO
      CLOSE(21)
      RU=SAV(:,:,:,1)
      RV=SAV(:,:,:,2)
      RW=SAV(:,:,:,3)
      RO=SAV(:,:,:,4)
      TT=SAV(:,:,:,5)
      BX=SAV(:,:,:,6)
      BY=SAV(:,:,:,7)
      BZ=SAV(:,:,:,8)
      DEALLOCATE(SAV)
C
      ND=1
      DO 100 N=1,NSTEPS
         CALL STEP
         IF (ND.LE.NDUMP) THEN
         IF (N.EQ.IDUMP(ND)) THEN
            WRITE(FNAME,'(A,I4.4,A,I4.4,A)') 'step',N,'_r',MYPE,'.bin'
            OPEN(21,FILE=FNAME,ACCESS='STREAM',FORM='UNFORMATTED',
     2           STATUS='REPLACE')
            WRITE(21) RU,RV,RW,RO,TT,BX,BY,BZ
            WRITE(21) DT,UMACH,TIMC,TIMT,DBLE(NIT)
            CLOSE(21)
            ND=ND+1
         ENDIF
         ENDIF
  100 CONTINUE
C
      CALL MPI_FINALIZE(IERR)
      STOP
      END
C-----------------------------------------------------------------------
C     Integer hash of (global index, component, seed) mapped to [-1/2,1/2).
      FUNCTION HASHR(IG,JG,KG,IC,ISEED)
      IMPLICIT NONE
      INTEGER IG,JG,KG,IC,ISEED
      REAL*8 HASHR
      INTEGER*8 H,P
      P=2147483647_8
      H=MOD(INT(IG,8)*7919_8+INT(JG,8)*104729_8+INT(KG,8)*1299709_8
     2     +INT(IC,8)*15485863_8+INT(ISEED,8)+12345_8,P)
      H=MOD(H*48271_8,P)
      H=MOD(H*48271_8,P)
      H=MOD(H*48271_8,P)
      HASHR=DBLE(H)/DBLE(P)-0.5D0
      RETURN
      END
