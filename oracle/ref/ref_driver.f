C======================================================================C
C  REF_DRIVER -- C-callable harness around the reference's own         C
C  two-electron routines (bertha-scf.f), compiled from the reference   C
C  sources where they lie.  This file is test infrastructure: it only  C
C  fills the COMMON blocks the reference routines read, calls them,    C
C  and copies their outputs out.  No reference routine is restated     C
C  here.  COMMON layouts are declared as the reference declares them   C
C  (bertha-scf.f:11533-11553) because Fortran COMMON is an ABI.        C
C                                                                      C
C  Entry points (BIND(C)):                                             C
C    ref_setup    basis/geometry -> FOCKIND, CARTIND, EQSAVE           C
C    ref_clmfast  DENT -> GDIR, GXCH via CLMFAST (bertha-scf.f:11489)  C
C    ref_erifast  one ERIFAST batch RR(MAXCD,16) (bertha-scf.f:12269)  C
C    ref_etable   one E0LL/E0SS block (bertha-scf.f:33378, 33511)      C
C    ref_funfm    FUNFM branch call (bertha-scf.f:32465)               C
C    ref_rmake    RMAKE batch (bertha-scf.f:31900)                     C
C    ref_labels   FOCKIND labels and LRGE offsets                      C
C======================================================================C
      SUBROUTINE REFSETUP(NCNTIN,XYZIN,LMXIN,NFNCIN,EXPIN,ISHP,
     &                    IHMLT,IRAC,ISYMT,NDIMO,NSKPO)
     &                    BIND(C,NAME='ref_setup')
      USE ISO_C_BINDING
      IMPLICIT DOUBLE PRECISION (A-H,O-Z)
      INCLUDE 'parameters.h'
      INCLUDE 'scfoptions.h'
      INTEGER(C_INT), VALUE :: NCNTIN,ISHP,IHMLT,IRAC,ISYMT
      INTEGER(C_INT) :: LMXIN(*),NFNCIN(MEL+1,*),NDIMO,NSKPO
      REAL(C_DOUBLE) :: XYZIN(3,*),EXPIN(MBS,MEL+1,*)
      COMMON/BDIM/NDIM,NSKP,NOCC,NVRT
      COMMON/BSET/BEXL(MBS,0:MEL,MCT),BXYZ(3,MCT),LRGE(MCT,MKP,MKP+1),
     &            KAPA(MKP,MCT),NFNC(0:MEL,MCT),NKAP(MCT),IQNC(MCT),NCNT
      COMMON/SHLL/ACFF,BCFF,FOPN,ICLS(MDM),IOPN(MDM),NCLS,NOPN,NOELEC
      DIMENSION ICDR11(MEQ,4),ICDI11(MEQ,4),ICDR21(MEQ,4),ICDI21(MEQ,4)
      DIMENSION ICDR11KM(4,MKP,MKP,MEQ,4),ICDI11KM(4,MKP,MKP,MEQ,4),
     &          ICDR21KM(4,MKP,MKP,MEQ,4),ICDI21KM(4,MKP,MKP,MEQ,4)
      COMMON/ICDC/ICDR11KM,ICDR11,ICDI11KM,ICDI11,
     &            ICDR21KM,ICDR21,ICDI21KM,ICDI21
      LOGICAL OPND
C
C     SILENCE THE REFERENCE'S UNIT 6/7 REPORTS
      INQUIRE(UNIT=7,OPENED=OPND)
      IF(.NOT.OPND) OPEN(UNIT=7,FILE='/dev/null',STATUS='UNKNOWN')
C     (UNIT 7 ALREADY HOLDS /dev/null, SO UNIT 6 TAKES /dev/zero)
      INQUIRE(FILE='/dev/zero',OPENED=OPND)
      IF(.NOT.OPND) OPEN(UNIT=6,FILE='/dev/zero',STATUS='UNKNOWN')
C
      CALL FACTRL
      CALL GAMGEN
      CALL CARTIND
C
C     START FROM THE ZEROED E(CD) SPARSITY LISTS OF A FRESH PROCESS
      ICDR11 = 0
      ICDI11 = 0
      ICDR21 = 0
      ICDI21 = 0
      ICDR11KM = 0
      ICDI11KM = 0
      ICDR21KM = 0
      ICDI21KM = 0
C
      IF(IHMLT.EQ.2) THEN
        HMLT = 'DHFB'
      ELSE
        HMLT = 'DHFR'
      ENDIF
      TREE = 'HFSCF'
      IF(ISHP.EQ.2) THEN
        SHAPE = 'ATOMIC'
      ELSEIF(ISHP.EQ.1) THEN
        SHAPE = 'DIATOM'
      ELSE
        SHAPE = 'NOSYMM'
      ENDIF
C
      RACAH1 = (IRAC.NE.0)
      INTSYM = (ISYMT.NE.0)
      SSSSI3 = .FALSE.
      SSSSI4 = .FALSE.
      GAUNT2 = .FALSE.
      VSDNSB = .FALSE.
      VSDNSC = .FALSE.
      DIISQC = .TRUE.
      DAMPFC = .FALSE.
      SCHWRZ = .FALSE.
      EQFILE = .TRUE.
      RCFILE = .TRUE.
      MQNSLC = .FALSE.
C
      NCNT = NCNTIN
      NDIM = 0
      DO IZ=1,NCNT
        DO J=1,3
          BXYZ(J,IZ) = XYZIN(J,IZ)
        ENDDO
        LMAX = LMXIN(IZ)
        NKAP(IZ) = 2*LMAX+1
        DO L=0,LMAX
          NFNC(L,IZ) = NFNCIN(L+1,IZ)
          DO IB=1,NFNC(L,IZ)
            BEXL(IB,L,IZ) = EXPIN(IB,L+1,IZ)
          ENDDO
          NDIM = NDIM + 4*(2*L+1)*NFNC(L,IZ)
          IF(L.NE.0) KAPA(2*L,IZ) = L
          KAPA(2*L+1,IZ) = -L-1
        ENDDO
      ENDDO
      NSKP = NDIM/2
      NOPN = 0
      ACFF = 0.0D0
      BCFF = 0.0D0
C
      CALL FOCKIND
      CALL EQSAVE
C
      NDIMO = NDIM
      NSKPO = NSKP
      RETURN
      END
C
C
      SUBROUTINE REFLABELS(ICN,KQN,MQN,LRGEO) BIND(C,NAME='ref_labels')
      USE ISO_C_BINDING
      IMPLICIT DOUBLE PRECISION (A-H,O-Z)
      INCLUDE 'parameters.h'
      INTEGER(C_INT) :: ICN(*),KQN(*),MQN(*),LRGEO(MCT,MKP,MKP+1)
      COMMON/BDIM/NDIM,NSKP,NOCC,NVRT
      COMMON/BSET/BEXL(MBS,0:MEL,MCT),BXYZ(3,MCT),LRGE(MCT,MKP,MKP+1),
     &            KAPA(MKP,MCT),NFNC(0:MEL,MCT),NKAP(MCT),IQNC(MCT),NCNT
      COMMON/QNMS/LABICN(MDM),LABKQN(MDM),LABMQN(MDM)
      DO I=1,NSKP
        ICN(I) = LABICN(I)
        KQN(I) = LABKQN(I)
        MQN(I) = LABMQN(I)
      ENDDO
      DO I=1,MCT
        DO J=1,MKP
          DO K=1,MKP+1
            LRGEO(I,J,K) = LRGE(I,J,K)
          ENDDO
        ENDDO
      ENDDO
      RETURN
      END
C
C
      SUBROUTINE REFCLMFAST(ILEVIN,DRE,DIM,GDRE,GDIM,GXRE,GXIM)
     &                      BIND(C,NAME='ref_clmfast')
      USE ISO_C_BINDING
      IMPLICIT DOUBLE PRECISION (A-H,O-Z)
      INCLUDE 'parameters.h'
      INTEGER(C_INT), VALUE :: ILEVIN
      REAL(C_DOUBLE) :: DRE(*),DIM(*),GDRE(*),GDIM(*),GXRE(*),GXIM(*)
      COMPLEX*16 DENC(MDM,MDM),DENO(MDM,MDM),DENT(MDM,MDM)
      COMPLEX*16 FOCK(MDM,MDM),OVLP(MDM,MDM),HNUC(MDM,MDM),
     &           HKIN(MDM,MDM),GDIR(MDM,MDM),GXCH(MDM,MDM),
     &           BDIR(MDM,MDM),BXCH(MDM,MDM),VANM(MDM,MDM),
     &           VSLF(MDM,MDM),VUEH(MDM,MDM),VWKR(MDM,MDM),
     &           VKSB(MDM,MDM),QDIR(MDM,MDM),QXCH(MDM,MDM),
     &           WDIR(MDM,MDM),WXCH(MDM,MDM),CPLE(MDM,MDM)
      COMMON/BDIM/NDIM,NSKP,NOCC,NVRT
      COMMON/DENS/DENC,DENO,DENT
      COMMON/LSHF/SHLEV(4),SHLV,ILEV
      COMMON/MTRX/FOCK,OVLP,HNUC,HKIN,GDIR,GXCH,BDIR,BXCH,VANM,VSLF,
     &            VUEH,VWKR,VKSB,QDIR,QXCH,WDIR,WXCH,CPLE
C
C     DENSITY IN: ROW-FASTEST (FORTRAN ORDER), LEADING DIMENSION NDIM
      ILEV = ILEVIN
      DO J=1,NDIM
        DO I=1,NDIM
          K = (J-1)*NDIM+I
          DENT(I,J) = DCMPLX(DRE(K),DIM(K))
          DENO(I,J) = DCMPLX(0.0D0,0.0D0)
          DENC(I,J) = DENT(I,J)
        ENDDO
      ENDDO
      CALL CLMFAST
      DO J=1,NDIM
        DO I=1,NDIM
          K = (J-1)*NDIM+I
          GDRE(K) = DREAL(GDIR(I,J))
          GDIM(K) = DIMAG(GDIR(I,J))
          GXRE(K) = DREAL(GXCH(I,J))
          GXIM(K) = DIMAG(GXCH(I,J))
        ENDDO
      ENDDO
      RETURN
      END
C
C
C     BRTFAST: CLOSED-SHELL BREIT MATRICES (BDIR, BXCH) FOR DENSITY DENT
      SUBROUTINE REFBRTFAST(ILEVIN,DRE,DIM,BDRE,BDIM,BXRE,BXIM)
     &                      BIND(C,NAME='ref_brtfast')
      USE ISO_C_BINDING
      IMPLICIT DOUBLE PRECISION (A-H,O-Z)
      INCLUDE 'parameters.h'
      INTEGER(C_INT), VALUE :: ILEVIN
      REAL(C_DOUBLE) :: DRE(*),DIM(*),BDRE(*),BDIM(*),BXRE(*),BXIM(*)
      COMPLEX*16 DENC(MDM,MDM),DENO(MDM,MDM),DENT(MDM,MDM)
      COMPLEX*16 FOCK(MDM,MDM),OVLP(MDM,MDM),HNUC(MDM,MDM),
     &           HKIN(MDM,MDM),GDIR(MDM,MDM),GXCH(MDM,MDM),
     &           BDIR(MDM,MDM),BXCH(MDM,MDM),VANM(MDM,MDM),
     &           VSLF(MDM,MDM),VUEH(MDM,MDM),VWKR(MDM,MDM),
     &           VKSB(MDM,MDM),QDIR(MDM,MDM),QXCH(MDM,MDM),
     &           WDIR(MDM,MDM),WXCH(MDM,MDM),CPLE(MDM,MDM)
      COMMON/BDIM/NDIM,NSKP,NOCC,NVRT
      COMMON/DENS/DENC,DENO,DENT
      COMMON/LSHF/SHLEV(4),SHLV,ILEV
      COMMON/MTRX/FOCK,OVLP,HNUC,HKIN,GDIR,GXCH,BDIR,BXCH,VANM,VSLF,
     &            VUEH,VWKR,VKSB,QDIR,QXCH,WDIR,WXCH,CPLE
      ILEV = ILEVIN
      DO J=1,NDIM
        DO I=1,NDIM
          K = (J-1)*NDIM+I
          DENT(I,J) = DCMPLX(DRE(K),DIM(K))
          DENO(I,J) = DCMPLX(0.0D0,0.0D0)
          DENC(I,J) = DENT(I,J)
        ENDDO
      ENDDO
      CALL BRTFAST
      DO J=1,NDIM
        DO I=1,NDIM
          K = (J-1)*NDIM+I
          BDRE(K) = DREAL(BDIR(I,J))
          BDIM(K) = DIMAG(BDIR(I,J))
          BXRE(K) = DREAL(BXCH(I,J))
          BXIM(K) = DIMAG(BXCH(I,J))
        ENDDO
      ENDDO
      RETURN
      END
C
C
C     OVRLP/ONEEL: OVLP, HNUC, HKIN FOR NUCLEAR CHARGES ZIN (MASS NUMBERS
C     AIN) WITH POINT (IMDL=0) OR GAUSSIAN (IMDL=1) NUCLEI; THE GAUSSIAN
C     WIDTH FOLLOWS THE REFERENCE'S RMS-RADIUS FORMULA (bertha-scf.f:
C     1182-1204) WITH NNUC = 0.
      SUBROUTINE REFONEEL(ZIN,AIN,IMDL,OVRE,OVIM,HNRE,HNIM,HKRE,HKIM)
     &                    BIND(C,NAME='ref_oneel')
      USE ISO_C_BINDING
      IMPLICIT DOUBLE PRECISION (A-H,O-Z)
      INCLUDE 'parameters.h'
      INCLUDE 'scfoptions.h'
      REAL(C_DOUBLE) :: ZIN(*),AIN(*),OVRE(*),OVIM(*),HNRE(*),HNIM(*),
     &                  HKRE(*),HKIM(*)
      INTEGER(C_INT), VALUE :: IMDL
      CHARACTER*5 NMDL
      COMPLEX*16 FOCK(MDM,MDM),OVLP(MDM,MDM),HNUC(MDM,MDM),
     &           HKIN(MDM,MDM),GDIR(MDM,MDM),GXCH(MDM,MDM),
     &           BDIR(MDM,MDM),BXCH(MDM,MDM),VANM(MDM,MDM),
     &           VSLF(MDM,MDM),VUEH(MDM,MDM),VWKR(MDM,MDM),
     &           VKSB(MDM,MDM),QDIR(MDM,MDM),QXCH(MDM,MDM),
     &           WDIR(MDM,MDM),WXCH(MDM,MDM),CPLE(MDM,MDM)
      COMMON/BDIM/NDIM,NSKP,NOCC,NVRT
      COMMON/BNUC/ZNUC(MCT),ANUC(MCT),TFMI(MCT),AFMI(MCT),RNUC(MCT),
     &            FNUC(MCT,0:MFT),XNUC(MCT,0:MFT),NNUC(MCT),NMDL(MCT)
      COMMON/BSET/BEXL(MBS,0:MEL,MCT),BXYZ(3,MCT),LRGE(MCT,MKP,MKP+1),
     &            KAPA(MKP,MCT),NFNC(0:MEL,MCT),NKAP(MCT),IQNC(MCT),NCNT
      COMMON/CONV/CHZ,CEV,CCM,CFM,CNG,CDB
      COMMON/PHYS/CV,EMSS,UMSS,TMSS,PMSS,PRAD,CMPW,GFREE,GFRMI,WEIN
      COMMON/MTRX/FOCK,OVLP,HNUC,HKIN,GDIR,GXCH,BDIR,BXCH,VANM,VSLF,
     &            VUEH,VWKR,VKSB,QDIR,QXCH,WDIR,WXCH,CPLE
      DO IZ=1,NCNT
        ZNUC(IZ) = ZIN(IZ)
        ANUC(IZ) = AIN(IZ)
        IF(IMDL.EQ.0) THEN
          NMDL(IZ) = 'POINT'
        ELSE
          NMDL(IZ) = 'GAUSS'
        ENDIF
        IF(INT(ZNUC(IZ)).EQ.1) THEN
          RFM = PRAD
        ELSEIF(ZNUC(IZ).GT.1.0D0.AND.ZNUC(IZ).LT.90.0D0) THEN
          RFM = 0.836D0*(ANUC(IZ)**(1.0D0/3.0D0)) + 0.57D0
        ELSE
          RFM = 0.770D0*(ANUC(IZ)**(1.0D0/3.0D0)) + 0.98D0
        ENDIF
        RNUC(IZ) = RFM/CFM
        NNUC(IZ)   = 0
        FNUC(IZ,0) = 1.0D0
        XNUC(IZ,0) = 1.5D0/(RNUC(IZ)*RNUC(IZ))
      ENDDO
      CALL ONEEL
      DO J=1,NDIM
        DO I=1,NDIM
          K = (J-1)*NDIM+I
          OVRE(K) = DREAL(OVLP(I,J))
          OVIM(K) = DIMAG(OVLP(I,J))
          HNRE(K) = DREAL(HNUC(I,J))
          HNIM(K) = DIMAG(HNUC(I,J))
          HKRE(K) = DREAL(HKIN(I,J))
          HKIM(K) = DIMAG(HKIN(I,J))
        ENDDO
      ENDDO
      RETURN
      END
C
C
C     CLMDIAG (SCHWARZ DIAGONAL BOUNDS) -> GDSC, ROW-FASTEST, LD = NDIM.
C     CLMDIAG SWITCHES RCFILE OFF FOR ITS OWN ERI CALLS; IT IS SWITCHED
C     BACK ON SO LATER ERIFAST CALLS RUN AS IN THE SCF.
      SUBROUTINE REFCLMDIAG(GOUT) BIND(C,NAME='ref_clmdiag')
      USE ISO_C_BINDING
      IMPLICIT DOUBLE PRECISION (A-H,O-Z)
      INCLUDE 'parameters.h'
      INCLUDE 'scfoptions.h'
      REAL(C_DOUBLE) :: GOUT(*)
      COMMON/BDIM/NDIM,NSKP,NOCC,NVRT
      COMMON/SWRZ/GDSC(MDM,MDM),BDSC(MDM,MDM)
      CALL CLMDIAG
      RCFILE = .TRUE.
      DO J=1,NDIM
        DO I=1,NDIM
          GOUT((J-1)*NDIM+I) = GDSC(I,J)
        ENDDO
      ENDDO
      RETURN
      END
C
C
C     SCHWRZ TOGGLE (HFSCF HARD-CODES .FALSE.)
      SUBROUTINE REFSCHWARZ(ION) BIND(C,NAME='ref_schwarz')
      USE ISO_C_BINDING
      IMPLICIT DOUBLE PRECISION (A-H,O-Z)
      INCLUDE 'parameters.h'
      INCLUDE 'scfoptions.h'
      INTEGER(C_INT), VALUE :: ION
      SCHWRZ = ION.NE.0
      RETURN
      END
C
C
      SUBROUTINE REFERIFAST(ICNTIN,KIN,MIN,IBASIN,JBASIN,ITNIN,ILEVIN,
     &                      RRRE,RRIM,MAXCDO) BIND(C,NAME='ref_erifast')
C     KIN(N)  = KAPPA INDEX (1..NKAP) OF BLOCK N
C     MIN(N)  = |MQN| INDEX (1..|KAPPA|) OF BLOCK N
C     REPLAYS THE PER-CLASS AND PER-BLOCK SET-UP THAT CLMFAST DOES
C     BEFORE ITS ERIFAST CALL (bertha-scf.f:11731-11745, 11800-11836,
C     11905-11910, 12091-12115) FOR ONE (IBAS,JBAS) AND ONE ITN PAIR.
      USE ISO_C_BINDING
      IMPLICIT DOUBLE PRECISION (A-H,O-Z)
      INCLUDE 'parameters.h'
      INCLUDE 'scfoptions.h'
      INTEGER(C_INT) :: ICNTIN(4),KIN(4),MIN(4),ITNIN(2),MAXCDO
      INTEGER(C_INT), VALUE :: IBASIN,JBASIN,ILEVIN
      REAL(C_DOUBLE) :: RRRE(*),RRIM(*)
      DIMENSION EXL(MBS,4),XYZ(3,4),ICNT(4),LQN(4),KQN(4),MQN(4),ITN(2)
      DIMENSION NBASL(4),MQNL(4)
      COMPLEX*16 RR(MB2,16)
      COMMON/BSET/BEXL(MBS,0:MEL,MCT),BXYZ(3,MCT),LRGE(MCT,MKP,MKP+1),
     &            KAPA(MKP,MCT),NFNC(0:MEL,MCT),NKAP(MCT),IQNC(MCT),NCNT
      COMMON/E0LL/E0LLFL(MFL,4),IAD0LL(MCT,MCT,MKP,MKP,MKP,MKP)
      COMMON/E0SS/E0SSFL(MFL,4),IAD0SS(MCT,MCT,MKP,MKP,MKP,MKP)
      COMMON/I2EL/PAB1,PAB2,PCD1,PCD2,NA1,NB1,NC1,ND1,NA2,NB2,NC2,ND2,
     &            IBAS,JBAS,MCNT,NADDAB,NADDCD,NBASL,MQNL,IQL,IQR
      COMMON/IQTT/IABLL,ICDLL,IABSS,ICDSS,IABLS,ICDLS,IABSL,ICDSL
      COMMON/IRCM/IEAB,IECD,NCD,IGAB,IRIJ(MBS,MBS)
      COMMON/ISCR/IMTX(MB2,11),ISCR(MB2),IMAP(MB2),IBCH,ITOG,MAXN
      COMMON/LSHF/SHLEV(4),SHLV,ILEV
      COMMON/SWRZ/GDSC(MDM,MDM),BDSC(MDM,MDM)
C
      ILEV = ILEVIN
      DO N=1,4
        ICNT(N) = ICNTIN(N)
        KQN(N)  = KAPA(KIN(N),ICNT(N))
        LQN(N)  = LVAL(KQN(N))
        MQN(N)  = 2*MIN(N)-1
        NBASL(N) = NFNC(LQN(N),ICNT(N))
        MQNL(N)  = MQN(N)
        DO IB=1,NBASL(N)
          EXL(IB,N) = BEXL(IB,LQN(N),ICNT(N))
        ENDDO
        DO J=1,3
          XYZ(J,N) = BXYZ(J,ICNT(N))
        ENDDO
      ENDDO
      ITN(1) = ITNIN(1)
      ITN(2) = ITNIN(2)
C
      DO IB=1,NBASL(1)
        DO JB=1,NBASL(2)
          IRIJ(IB,JB) = 1
        ENDDO
      ENDDO
      CALL E0LLCD0(ICNT,LQN,NBASL)
      IF(ILEV.NE.1) CALL E0SSCD0(ICNT,LQN,NBASL)
C
      IABLL = IAD0LL(ICNT(1),ICNT(2),KIN(1),KIN(2),MIN(1),MIN(2))
      IABSS = IAD0SS(ICNT(1),ICNT(2),KIN(1),KIN(2),MIN(1),MIN(2))
      ICDLL = IAD0LL(ICNT(3),ICNT(4),KIN(3),KIN(4),MIN(3),MIN(4))
      ICDSS = IAD0SS(ICNT(3),ICNT(4),KIN(3),KIN(4),MIN(3),MIN(4))
      IEAB = 1
      IECD = 1
      IGAB = 1
      NC = KIN(3)-2*LQN(3)
      ND = KIN(4)-2*LQN(4)
      NCD = 2*NC+ND+1
      IBAS = IBASIN
      JBAS = JBASIN
      ITOG = 0
      CALL SCHWARZ(GDSC,1.0D-12)
C
      CALL ERIFAST(RR,XYZ,ICNT,KQN,MQN,NBASL,EXL,IBAS,JBAS,ITN)
C
      MAXCDO = NBASL(3)*NBASL(4)
      DO ITG=1,16
        DO M=1,MAXCDO
          K = (ITG-1)*MAXCDO+M
          RRRE(K) = DREAL(RR(M,ITG))
          RRIM(K) = DIMAG(RR(M,ITG))
        ENDDO
      ENDDO
      RETURN
      END
C
C
      SUBROUTINE REFBIIFAST(ICNTIN,KIN,MIN,IBASIN,JBASIN,IGNT,
     &                      RRRE,RRIM,MAXCDO) BIND(C,NAME='ref_biifast')
C     REPLAYS THE PER-CLASS AND PER-BLOCK SET-UP THAT BRTFAST DOES BEFORE
C     ITS BIIFAST CALL (bertha-scf.f:17396-17410, 17420-17424, 17517-17521,
C     17560-17590) FOR ONE (IBAS,JBAS) OF A FRESH CLASS.  IGNT = GAUNT2.
      USE ISO_C_BINDING
      IMPLICIT DOUBLE PRECISION (A-H,O-Z)
      INCLUDE 'parameters.h'
      INCLUDE 'scfoptions.h'
      INTEGER(C_INT) :: ICNTIN(4),KIN(4),MIN(4),MAXCDO
      INTEGER(C_INT), VALUE :: IBASIN,JBASIN,IGNT
      REAL(C_DOUBLE) :: RRRE(*),RRIM(*)
      DIMENSION EXL(MBS,4),XYZ(3,4),ICNT(4),LQN(4),KQN(4),MQN(4),ITN(2)
      DIMENSION NBASL(4),MQNL(4)
      COMPLEX*16 RR(MB2,16)
      COMMON/BSET/BEXL(MBS,0:MEL,MCT),BXYZ(3,MCT),LRGE(MCT,MKP,MKP+1),
     &            KAPA(MKP,MCT),NFNC(0:MEL,MCT),NKAP(MCT),IQNC(MCT),NCNT
      COMMON/EILS/EILSFL(MFL,12),IADILS(MCT,MCT,MKP,MKP,MKP,MKP)
      COMMON/I2EL/PAB1,PAB2,PCD1,PCD2,NA1,NB1,NC1,ND1,NA2,NB2,NC2,ND2,
     &            IBAS,JBAS,MCNT,NADDAB,NADDCD,NBASL,MQNL,IQL,IQR
      COMMON/IQTT/IABLL,ICDLL,IABSS,ICDSS,IABLS,ICDLS,IABSL,ICDSL
      COMMON/IRCM/IEAB,IECD,NCD,IGAB,IRIJ(MBS,MBS)
      COMMON/ISCR/IMTX(MB2,11),ISCR(MB2),IMAP(MB2),IBCH,ITOG,MAXN
      COMMON/SWRZ/GDSC(MDM,MDM),BDSC(MDM,MDM)
C
      BGAUNT = IGNT.NE.0
      DO N=1,4
        ICNT(N) = ICNTIN(N)
        KQN(N)  = KAPA(KIN(N),ICNT(N))
        LQN(N)  = LVAL(KQN(N))
        MQN(N)  = 2*MIN(N)-1
        NBASL(N) = NFNC(LQN(N),ICNT(N))
        MQNL(N)  = MQN(N)
        DO IB=1,NBASL(N)
          EXL(IB,N) = BEXL(IB,LQN(N),ICNT(N))
        ENDDO
        DO J=1,3
          XYZ(J,N) = BXYZ(J,ICNT(N))
        ENDDO
      ENDDO
      ITN(1) = 2
      ITN(2) = 2
      DO IB=1,NBASL(1)
        DO JB=1,NBASL(2)
          IRIJ(IB,JB) = 1
        ENDDO
      ENDDO
      CALL EILSCD0(ICNT,LQN,NBASL)
      IABLS = IADILS(ICNT(1),ICNT(2),KIN(1),KIN(2),MIN(1),MIN(2))
      ICDLS = IADILS(ICNT(3),ICNT(4),KIN(3),KIN(4),MIN(3),MIN(4))
      IEAB = 1
      IECD = 1
      IGAB = 1
      NC = KIN(3)-2*LQN(3)
      ND = KIN(4)-2*LQN(4)
      NCD = 2*NC+ND+1
      IBAS = IBASIN
      JBAS = JBASIN
      ITOG = 0
      CALL SCHWARZ(BDSC,1.0D-12)
C
      CALL BIIFAST(RR,XYZ,ICNT,KQN,MQN,NBASL,EXL,IBAS,JBAS,ITN)
C
      MAXCDO = NBASL(3)*NBASL(4)
      DO ITG=1,16
        DO M=1,MAXCDO
          K = (ITG-1)*MAXCDO+M
          RRRE(K) = DREAL(RR(M,ITG))
          RRIM(K) = DIMAG(RR(M,ITG))
        ENDDO
      ENDDO
      RETURN
      END
C
C
      SUBROUTINE REFETABLE(ITT,ICA,ICB,KA,KB,MA,MB,ENT,EOUT,NTUVO,
     &                     MAXABO) BIND(C,NAME='ref_etable')
C     ITT = 1 (E0LL), 4 (E0SS) OR 2 (EILS, 12 COMPONENTS);
C     EOUT(NCOMP*NTUV*MAXAB) IN THE STORED ORDER [COMPONENT][TUV][IJ]
C     WITH IJ FASTEST (bertha-scf.f:33493, 33757).
      USE ISO_C_BINDING
      IMPLICIT DOUBLE PRECISION (A-H,O-Z)
      INCLUDE 'parameters.h'
      INTEGER(C_INT), VALUE :: ITT,ICA,ICB,KA,KB,MA,MB,ENT
      INTEGER(C_INT) :: NTUVO,MAXABO
      REAL(C_DOUBLE) :: EOUT(*)
      COMMON/BSET/BEXL(MBS,0:MEL,MCT),BXYZ(3,MCT),LRGE(MCT,MKP,MKP+1),
     &            KAPA(MKP,MCT),NFNC(0:MEL,MCT),NKAP(MCT),IQNC(MCT),NCNT
      COMMON/E0LL/E0LLFL(MFL,4),IAD0LL(MCT,MCT,MKP,MKP,MKP,MKP)
      COMMON/E0SS/E0SSFL(MFL,4),IAD0SS(MCT,MCT,MKP,MKP,MKP,MKP)
      COMMON/EILS/EILSFL(MFL,12),IADILS(MCT,MCT,MKP,MKP,MKP,MKP)
      LA = LVAL(KAPA(KA,ICA))
      LB = LVAL(KAPA(KB,ICB))
      MAXABO = NFNC(LA,ICA)*NFNC(LB,ICB)
      IF(ITT.EQ.2) THEN
        LAM = LA+LB+1
        IAD = IADILS(ICA,ICB,KA,KB,MA,MB)
        NTUVO = (LAM+1)*(LAM+2)*(LAM+3)/6
        N = NTUVO*MAXABO
        IF(ENT.LT.12*N) THEN
          NTUVO = -1
          RETURN
        ENDIF
        DO IC=1,12
          DO K=1,N
            EOUT((IC-1)*N+K) = EILSFL(IAD+K,IC)
          ENDDO
        ENDDO
        RETURN
      ENDIF
      IF(ITT.EQ.1) THEN
        LAM = LA+LB
        IAD = IAD0LL(ICA,ICB,KA,KB,MA,MB)
      ELSE
        LAM = LA+LB+2
        IAD = IAD0SS(ICA,ICB,KA,KB,MA,MB)
      ENDIF
      NTUVO = (LAM+1)*(LAM+2)*(LAM+3)/6
      N = NTUVO*MAXABO
      IF(ENT.LT.4*N) THEN
        NTUVO = -1
        RETURN
      ENDIF
      DO IC=1,4
        DO K=1,N
          IF(ITT.EQ.1) THEN
            EOUT((IC-1)*N+K) = E0LLFL(IAD+K,IC)
          ELSE
            EOUT((IC-1)*N+K) = E0SSFL(IAD+K,IC)
          ENDIF
        ENDDO
      ENDDO
      RETURN
      END
C
C
      SUBROUTINE REFFUNFM(TIN,NIN,LAMIN,ITYPE,FOUT)
     &                    BIND(C,NAME='ref_funfm')
C     FOUT(N,LAM+1) FOR ONE EXPLICIT FUNFM BRANCH
      USE ISO_C_BINDING
      IMPLICIT DOUBLE PRECISION (A-H,O-Z)
      INCLUDE 'parameters.h'
      INTEGER(C_INT), VALUE :: NIN,LAMIN,ITYPE
      REAL(C_DOUBLE) :: TIN(*),FOUT(*)
      DIMENSION FM(MB2,ML4),T(MB2)
      DO M=1,NIN
        T(M) = TIN(M)
      ENDDO
      CALL FUNFM(FM,T,NIN,LAMIN,ITYPE)
      DO K=1,LAMIN+1
        DO M=1,NIN
          FOUT((K-1)*NIN+M) = FM(M,K)
        ENDDO
      ENDDO
      RETURN
      END
C
C
      SUBROUTINE REFRMAKE(QPIN,APHIN,NIN,LAMIN,RCOUT)
     &                    BIND(C,NAME='ref_rmake')
C     RCOUT(N,NTUV) IN CARTIND ORDER; QPIN(N,3) = Q-P
      USE ISO_C_BINDING
      IMPLICIT DOUBLE PRECISION (A-H,O-Z)
      INCLUDE 'parameters.h'
      INTEGER(C_INT), VALUE :: NIN,LAMIN
      REAL(C_DOUBLE) :: QPIN(*),APHIN(*),RCOUT(*)
      DIMENSION QP(MB2,3),APH(MB2),RC(MB2,MRC)
      DO M=1,NIN
        APH(M) = APHIN(M)
        DO J=1,3
          QP(M,J) = QPIN((J-1)*NIN+M)
        ENDDO
      ENDDO
      CALL RMAKE(RC,QP,APH,NIN,LAMIN)
      NTUV = (LAMIN+1)*(LAMIN+2)*(LAMIN+3)/6
      DO K=1,NTUV
        DO M=1,NIN
          RCOUT((K-1)*NIN+M) = RC(M,K)
        ENDDO
      ENDDO
      RETURN
      END
