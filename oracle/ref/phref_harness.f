c-----------------------------------------------------------------------
c  Test harness around the reference's own compiled routines (built by
c  oracle/ref/Makefile from the reference sources into oracle/_ref/).
c  It fills the COMMON blocks of common.h and the element blocks of
c  module pointer_data the way readnblk/genblk/genbkb/genint/genshp
c  would for a linear-tet mesh, and exposes ElmGMRs, e3, SparseAp,
c  Spsi3pre, bc3per, i3LU, itrPredict, itrCorrect and itrBC, and the
c  matrix-free solver's ElmMFG, ItrRes, itrFDI, Au1MFG, Au2MFG and
c  yshuffle, to C through bind(C) wrappers.  SolMFG itself is not built
c  either: it calls rstat with two arguments where rstat takes three
c  (solmfg.f:368, rstat.f:1), so its last array is undefined.  SolGMRs itself is not built: its history print
c  (rstat.f:115) needs TMRC from tmrc.c, whose header includes the
c  CMake-generated FCMangle.h.
c  Test infrastructure only: the product never links or calls this.
c-----------------------------------------------------------------------
      module phref_data
        use iso_c_binding
        real*8, allocatable :: hx(:,:), hBC(:,:), hrerr(:,:)
        integer, allocatable :: hiBC(:), hiper(:), hcolm(:), hrowp(:)
        integer :: hilwork(1)
        real*8 :: hshp(6,32,125), hshgl(6,3,32,125)
        real*8 :: hshpb(6,32,125), hshglb(6,3,32,125)
      end module phref_data

      subroutine phref_setup(nshg_, nel_, nelb_, nnz_, ndofBC_, ibk,
     &                       x_, ien_, ienb_, iBC_, BC_, iper_,
     &                       colm_, rowp_, dprm, iprm)
     &           bind(C, name='phref_setup')
      use iso_c_binding
      use pointer_data
      use phref_data
      include "common.h"
      integer(c_long_long), value :: nshg_, nel_, nelb_, nnz_
      integer(c_int), value :: ndofBC_, ibk
      real(c_double) :: x_(nshg_,3), BC_(nshg_,ndofBC_), dprm(14)
      integer(c_int) :: ien_(nel_,4), ienb_(nelb_,4), iBC_(nshg_),
     &                  iper_(nshg_), colm_(nshg_+1), rowp_(nnz_),
     &                  iprm(9)
      integer :: iblk, iel0, nprob, k, q, a, j
      real*8 :: qa4, qb4, qA, qB, qC, rr(4,3), r3(3,3), el(4)
      interface
        subroutine initphcommonvars() bind(C, name='initPhastaCommonVars')
        end subroutine
      end interface
c.... the reference's own defaults for the COMMON blocks (common.f)
      call initphcommonvars()
c.... sizes (/conpar/, /workfc/, /newdim/)
      numpe = 1
      myrank = 0
      master = 0
      nshg = nshg_
      numnp = nshg_
      nshgt = nshg_
      numel = nel_
      numelb = nelb_
      nnz_tot = nnz_
      nnz = 35
      ndof = 5
      nflow = 5
      nen = 4
      nedof = 20
      navier = 1
      ndofBC = ndofBC_
      ndiBCB = 2
      ndBCB = 6
      ipord = 1
      nshape = 4
      nshapeb = 3
      ibksiz = ibk
c.... physics (/mmatpar/, /matdat/, /genpar/, /outpar/, /timdat/)
      Rgas = dprm(1)
      gamma = dprm(2)
      gamma1 = dprm(3)
      pr = dprm(4)
      datmat(1,2,1) = dprm(5)
      dtsfct = dprm(6)
      taucfct = dprm(7)
      temper = dprm(8)
      almi = dprm(9)
      alfi = dprm(10)
      gami = dprm(11)
      Dtgl = dprm(12)
      etol = dprm(13)
      epstol(1) = dprm(13)
      epsM = dprm(14)
      Kspace = iprm(1)
      nGMRES = iprm(2)
      minIters = iprm(3)
      lhs = iprm(4)
      iprec = iprm(5)
      LHSupd(1) = 1
      ipred = iprm(6)
      idiff = iprm(7)
      iDC = iprm(8)
      itau = iprm(9)
      idflx = 0
      if (idiff >= 1) idflx = (nflow-1) * nsd
      isurf = 0
      rhoinf(1) = -1.0d0
      Jactyp = 0
c.... Weak Form: SUPG, the input default (input.config:226,
c     input_fform.cc:645-648); read only by e3bdg
      ivart = 2
      iabres = 0
      itseq = 1
      iter = 1
      nitr = 0
      lstep = 0
      loctim(1) = 0
c.... history output of rstat goes to a discarded unit
      ihist = 77
      open(unit=77, file='/dev/null')
      ierror = 78
      ferror = 'phref_error.log'
      title = 'phref harness'
c.... quadrature and shapes of the linear tet (symtet.c, symtri.c,
c     genint.f:74-75, genshp.f:34-35, genshpb.f:33-34)
      qa4 = 0.5854101966249685d0
      qb4 = 0.1381966011250150d0
      rr(1,:) = (/ qa4, qb4, qb4 /)
      rr(2,:) = (/ qb4, qa4, qb4 /)
      rr(3,:) = (/ qb4, qb4, qa4 /)
      rr(4,:) = (/ qb4, qb4, qb4 /)
      qA = 0.3333333333333330d0
      qB = 0.666666666666667d0
      qC = 0.166666666666667d0
      r3(1,:) = (/ qB, qC, qC /)
      r3(2,:) = (/ qC, qB, qC /)
      r3(3,:) = (/ qC, qC, qB /)
      nint(1) = 4
      nintb(1) = 3
      hshp = 0
      hshgl = 0
      hshpb = 0
      hshglb = 0
      do q = 1, 4
         el(1) = rr(q,1)
         el(2) = rr(q,2)
         el(3) = rr(q,3)
         el(4) = 1.0d0 - rr(q,1) - rr(q,2) - rr(q,3)
         do a = 1, 4
            hshp(1,a,q) = el(a)
         enddo
         Qwt(1,q) = (4.0d0/3.0d0) * 0.25d0
      enddo
      do q = 1, 3
         el(1) = r3(q,1)
         el(2) = r3(q,2)
         el(3) = r3(q,3)
         el(4) = 1.0d0 - r3(q,1) - r3(q,2) - r3(q,3)
         do a = 1, 4
            hshpb(1,a,q) = el(a)
         enddo
         Qwtb(1,q) = 2.0d0 * qA
      enddo
      do q = 1, 4
         do j = 1, 3
            hshgl(1,j,j,q) = 0.5d0
            hshgl(1,j,4,q) = -0.5d0
         enddo
      enddo
      do q = 1, 3
         do j = 1, 3
            hshglb(1,j,j,q) = 0.5d0
            hshglb(1,j,4,q) = -0.5d0
         enddo
      enddo
c.... mesh copies
      if (allocated(hx)) deallocate(hx, hBC, hiBC, hiper, hcolm,
     &                               hrowp, hrerr)
      allocate(hx(nshg_,3), hBC(nshg_,ndofBC_), hiBC(nshg_),
     &         hiper(nshg_), hcolm(nshg_+1), hrowp(nnz_),
     &         hrerr(nshg_,10))
      hx = x_
      hBC = BC_
      hiBC = iBC_
      hiper = iper_
      hcolm = colm_
      hrowp = rowp_
      hrerr = 0
c.... interior element blocks (genblk.f:111-160)
      nelblk = 0
      iel0 = 1
      do k = 1, int(nel_), ibk
         nelblk = nelblk + 1
         iblk = nelblk
         nprob = min(ibk, int(nel_) - k + 1)
         lcblk(1,iblk) = iel0
         lcblk(2,iblk) = 1
         lcblk(3,iblk) = 1
         lcblk(4,iblk) = 1
         lcblk(5,iblk) = 4
         lcblk(6,iblk) = 4
         lcblk(7,iblk) = 1
         lcblk(8,iblk) = 5
         lcblk(9,iblk) = 0
         lcblk(10,iblk) = 4
         allocate(mien(iblk)%p(nprob,4), mmat(iblk)%p(nprob),
     &            mxmudmi(iblk)%p(nprob,32))
         mien(iblk)%p(:,:) = ien_(k:k+nprob-1,:)
         mmat(iblk)%p(:) = 1
         mxmudmi(iblk)%p(:,:) = 0
         iel0 = iel0 + nprob
      enddo
      lcblk(1,nelblk+1) = iel0
c.... boundary element blocks (genbkb.f)
      nelblb = 0
      iel0 = 1
      do k = 1, int(nelb_), ibk
         nelblb = nelblb + 1
         iblk = nelblb
         nprob = min(ibk, int(nelb_) - k + 1)
         lcblkb(1,iblk) = iel0
         lcblkb(2,iblk) = 1
         lcblkb(3,iblk) = 1
         lcblkb(4,iblk) = 1
         lcblkb(5,iblk) = 4
         lcblkb(6,iblk) = 3
         lcblkb(7,iblk) = 1
         lcblkb(8,iblk) = 5
         lcblkb(9,iblk) = 4
         lcblkb(10,iblk) = 3
         allocate(mienb(iblk)%p(nprob,4), mmatb(iblk)%p(nprob),
     &            miBCB(iblk)%p(nprob,2), mBCB(iblk)%p(nprob,3,6))
         mienb(iblk)%p(:,:) = ienb_(k:k+nprob-1,:)
         mmatb(iblk)%p(:) = 1
         miBCB(iblk)%p(:,:) = 0
         mBCB(iblk)%p(:,:,:) = 0
         iel0 = iel0 + nprob
      enddo
      lcblkb(1,nelblb+1) = iel0
      end subroutine

      subroutine phref_elmgmrs(y, ac, res, rmes, BDiag, lhsK)
     &           bind(C, name='phref_elmgmrs')
      use iso_c_binding
      use phref_data
      include "common.h"
      real(c_double) :: y(nshg,5), ac(nshg,5), res(nshg,5),
     &                  rmes(nshg,5), BDiag(nshg,5,5), lhsK(25,nnz_tot)
      call ElmGMRs(y, ac, hx, hshp, hshgl, hiBC, hBC, hshpb, hshglb,
     &             res, rmes, BDiag, hiper, hilwork, lhsK, hcolm,
     &             hrowp, hrerr)
      end subroutine

      subroutine phref_sparseap(lhsK, p) bind(C, name='phref_sparseap')
      use iso_c_binding
      use phref_data
      include "common.h"
      real(c_double) :: lhsK(25,nnz_tot), p(nshg,5)
      call SparseAp(hiper, hilwork, hiBC, hcolm, hrowp, lhsK, p)
      end subroutine

      subroutine phref_spsi3pre(BDiag, lhsK)
     &           bind(C, name='phref_spsi3pre')
      use iso_c_binding
      use phref_data
      include "common.h"
      real(c_double) :: BDiag(nshg,5,5), lhsK(25,nnz_tot)
      call Spsi3pre(BDiag, lhsK, hcolm, hrowp)
      end subroutine

      subroutine phref_bc3per(res, nq) bind(C, name='phref_bc3per')
      use iso_c_binding
      use phref_data
      include "common.h"
      integer(c_int), value :: nq
      real(c_double) :: res(nshg,nq)
      call bc3per(hiBC, res, hiper, hilwork, nq)
      end subroutine

      subroutine phref_e3(yl, acl, xl, rl, EGmass)
     &           bind(C, name='phref_e3')
      use iso_c_binding
      use phref_data
      include "common.h"
      real(c_double) :: yl(1,4,5), acl(1,4,5), xl(1,4,3), rl(1,4,5),
     &                  EGmass(1,20,20)
      real*8 :: rml(1,4,5), xmudmi(1,4), BDiagl(1,4,5,5), ql(1,4,1),
     &          sgn(1,4), rlsl(1,4,6), rerrl(1,4,6), ytargetl(1,4,5)
      npro = 1
      nshl = 4
      nenl = 4
      ngauss = 4
      lcsyst = 1
      ires = 1
      rl = 0
      rml = 0
      BDiagl = 0
      EGmass = 0
      xmudmi = 0
      rlsl = 0
      call e3(yl, yl, acl, hshp(1,1:4,1:4), hshgl(1,:,1:4,1:4), xl,
     &        rl, rml, xmudmi, BDiagl, ql, sgn, rlsl, EGmass, rerrl,
     &        ytargetl)
      end subroutine

      subroutine phref_i3lu(Diag, r, kcode) bind(C, name='phref_i3lu')
      use iso_c_binding
      include "common.h"
      real(c_double) :: Diag(nshg,5,5), r(nshg,5)
      integer(c_int), value :: kcode
      character*8 code
      if (kcode .eq. 0) code = 'LU_Fact '
      if (kcode .eq. 1) code = 'forward '
      if (kcode .eq. 2) code = 'backward'
      if (kcode .eq. 3) code = 'product '
      call i3LU(Diag, r, code)
      end subroutine

c.... matrix-free GMRES pieces.  ItrRes sets iprec = 0 and ires = 2 in
c     COMMON (itrres.f:38-40); the wrappers restore both so that later
c     calls see the caller's flags.
      subroutine phref_elmmfg(y, ac, res, rmes, BDiag)
     &           bind(C, name='phref_elmmfg')
      use iso_c_binding
      use phref_data
      include "common.h"
      real(c_double) :: y(nshg,5), ac(nshg,5), res(nshg,5),
     &                  rmes(nshg,5), BDiag(nshg,5,5)
      integer :: isav(2)
      isav = (/ iprec, ires /)
      call ElmMFG(y, ac, hx, hshp, hshgl, hiBC, hBC, hshpb, hshglb,
     &            res, rmes, BDiag, hiper, hilwork, hrerr)
      iprec = isav(1)
      ires = isav(2)
      end subroutine

      subroutine phref_itrres(yp, yc, ac, rmes, iab)
     &           bind(C, name='phref_itrres')
      use iso_c_binding
      use phref_data
      include "common.h"
      real(c_double) :: yp(nshg,5), yc(nshg,5), ac(nshg,5),
     &                  rmes(nshg,5)
      integer(c_int), value :: iab
      integer :: isav(2)
      isav = (/ iprec, ires /)
      iabres = iab
      call ItrRes(yp, yc, hx, hshp, hshgl, hiBC, hBC, hshpb, hshglb,
     &            rmes, hiper, hilwork, ac)
      iabres = 0
      iprec = isav(1)
      ires = isav(2)
      end subroutine

      subroutine phref_itrfdi(ypre, y, ac, rmes, u, BDiag, eg)
     &           bind(C, name='phref_itrfdi')
      use iso_c_binding
      use phref_data
      include "common.h"
      real(c_double) :: ypre(nshg,5), y(nshg,5), ac(nshg,5),
     &                  rmes(nshg,5), u(nshg,5), BDiag(nshg,5,5), eg
      integer :: isav(2)
      isav = (/ iprec, ires /)
      ires = 3
      call itrFDI(ypre, y, ac, hx, rmes, u, BDiag, hiBC, hBC, hiper,
     &            hilwork, hshp, hshgl, hshpb, hshglb)
      eg = eGMRES
      iprec = isav(1)
      ires = isav(2)
      end subroutine

c.... SolMFG sets ires = 2 before its Krylov loop (solmfg.f:171), which
c     keeps itrBC from touching its second (aliased) argument
      subroutine phref_au1mfg(ypre, y, ac, rmes, res, u, BDiag, eg)
     &           bind(C, name='phref_au1mfg')
      use iso_c_binding
      use phref_data
      include "common.h"
      real(c_double) :: ypre(nshg,5), y(nshg,5), ac(nshg,5),
     &                  rmes(nshg,5), res(nshg,5), u(nshg,5),
     &                  BDiag(nshg,5,5)
      real(c_double), value :: eg
      integer :: isav(2)
      isav = (/ iprec, ires /)
      eGMRES = eg
      ires = 2
      call Au1MFG(ypre, y, ac, hx, rmes, res, u, BDiag, hiBC, hBC,
     &            hiper, hilwork, hshp, hshgl, hshpb, hshglb)
      iprec = isav(1)
      ires = isav(2)
      end subroutine

      subroutine phref_au2mfg(ypre, y, ac, rmes, res, Dy, u, BDiag)
     &           bind(C, name='phref_au2mfg')
      use iso_c_binding
      use phref_data
      include "common.h"
      real(c_double) :: ypre(nshg,5), y(nshg,5), ac(nshg,5),
     &                  rmes(nshg,5), res(nshg,5), Dy(nshg,5),
     &                  u(nshg,5), BDiag(nshg,5,5)
      integer :: isav(2)
      isav = (/ iprec, ires /)
      ires = 2
      call Au2MFG(ypre, y, ac, hx, rmes, res, Dy, u, BDiag, hiBC, hBC,
     &            hiper, hilwork, hshp, hshgl, hshpb, hshglb)
      iprec = isav(1)
      ires = isav(2)
      end subroutine

      subroutine phref_yshuffle(r, kcode) bind(C, name='phref_yshuffle')
      use iso_c_binding
      include "common.h"
      real(c_double) :: r(nshg,5)
      integer(c_int), value :: kcode
      character*8 code
      code = 'old2new '
      if (kcode .eq. 1) code = 'new2old '
      call yshuffle(r, code)
      end subroutine

      subroutine phref_predict_bc(yold, acold, y, ac)
     &           bind(C, name='phref_predict_bc')
      use iso_c_binding
      use phref_data
      include "common.h"
      real(c_double) :: yold(nshg,5), acold(nshg,5), y(nshg,5),
     &                  ac(nshg,5)
      call itrPredict(yold, acold, y, ac)
      call itrBC(y, ac, hiBC, hBC, hiper, hilwork)
      end subroutine

c.... itrSetup for spectral radius rinf and predictor ipin: returns
c     almi, alfi, gami, the predictor it settles on, and y after its
c     ipred = 4 start-up shift; the COMMON values it changes are restored
      subroutine phref_itrsetup(rinf, ipin, y, acold, fac, ipout)
     &           bind(C, name='phref_itrsetup')
      use iso_c_binding
      use phref_data
      include "common.h"
      real(c_double), value :: rinf
      integer(c_int), value :: ipin
      real(c_double) :: y(nshg,5), acold(nshg,5), fac(3)
      integer(c_int) :: ipout
      real*8 :: sav(7)
      integer :: isav(2)
      sav = (/ almi, alfi, gami, Dtgl, CFLfld, CFLsld, rhoinf(1) /)
      isav = (/ ipred, Jactyp /)
      rhoinf(1) = rinf
      ipred = ipin
      itseq = 1
      Delt(1) = one / Dtgl
      call itrSetup(y, acold)
      fac(1) = almi
      fac(2) = alfi
      fac(3) = gami
      ipout = ipred
      almi = sav(1)
      alfi = sav(2)
      gami = sav(3)
      Dtgl = sav(4)
      CFLfld = sav(5)
      CFLsld = sav(6)
      rhoinf(1) = sav(7)
      ipred = isav(1)
      Jactyp = isav(2)
      end subroutine

      subroutine phref_update_bc(yold, acold, y, ac)
     &           bind(C, name='phref_update_bc')
      use iso_c_binding
      use phref_data
      include "common.h"
      real(c_double) :: yold(nshg,5), acold(nshg,5), y(nshg,5),
     &                  ac(nshg,5)
      call itrUpdate(yold, acold, y, ac)
      call itrBC(yold, acold, hiBC, hBC, hiper, hilwork)
      end subroutine

      subroutine phref_correct_bc(y, ac, yold, acold, Dy)
     &           bind(C, name='phref_correct_bc')
      use iso_c_binding
      use phref_data
      include "common.h"
      real(c_double) :: yold(nshg,5), acold(nshg,5), y(nshg,5),
     &                  ac(nshg,5), Dy(nshg,5)
      call itrCorrect(y, ac, yold, acold, Dy)
      call itrBC(y, ac, hiBC, hBC, hiper, hilwork)
      end subroutine
