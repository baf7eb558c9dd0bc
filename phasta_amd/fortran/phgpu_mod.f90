!-----------------------------------------------------------------------
! module phgpu: iso_c_binding interfaces of libphgpu.so (include/phgpu.h)
! for a Fortran host.  The derived types mirror the C structs field by
! field; array arguments are the host's own Fortran arrays, passed by
! reference exactly as the reference passes them to SolGMRs/ElmGMRs
! (solgmr.f:368-433, elmgmr.f:280-323).
!-----------------------------------------------------------------------
module phgpu
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: PHG_FP64 = 0, PHG_MIXED = 1, PHG_FP64_MATRIX_FREE = 2

  type, bind(C) :: phg_params
    real(c_double) :: Rgas, gamma, gamma1, pr, mu
    real(c_double) :: dtsfct, taucfct, temper
    real(c_double) :: almi, alfi, gami, Dtgl
    real(c_double) :: etol, epsM
    integer(c_int) :: Kspace, nGMRES, minIters
    integer(c_int) :: lhs, iprec
    integer(c_int) :: ipred, idiff, iter, idc, itau
  end type

  type, bind(C) :: phg_mesh
    integer(c_long_long) :: nshg, nel, nelb, nnz_tot
    integer(c_int) :: ndofBC, reserved
    type(c_ptr) :: x, ien, ienb, iBC, BC, iper, colm, rowp
  end type

  type, bind(C) :: phg_stats
    integer(c_int) :: iKs, lGMRES, ntotGM, converged
    real(c_double) :: unorm0, epsnrm, echeck
    real(c_double) :: totres(2)
    real(c_double) :: reslast
    integer(c_long_long) :: nnan, nlarge
    real(c_double) :: ms_assembly, ms_solve, ms_total
    real(c_double) :: ms_spmv, ms_mgs
    integer(c_int) :: n_spmv, n_mgs
    integer(c_int) :: refine_cycles, pc_fallback
    real(c_double) :: orth_loss
    real(c_double) :: ms_point, ms_blocks
  end type

  interface
    integer(c_int) function phg_init(mesh, prm, precision, ctx) bind(C, name='phg_init')
      import :: c_int, c_ptr, phg_mesh, phg_params
      type(phg_mesh), intent(in) :: mesh
      type(phg_params), intent(in) :: prm
      integer(c_int), value :: precision
      type(c_ptr), intent(out) :: ctx
    end function

    integer(c_int) function phg_finalize(ctx) bind(C, name='phg_finalize')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function phg_set_params(ctx, prm) bind(C, name='phg_set_params')
      import :: c_int, c_ptr, phg_params
      type(c_ptr), value :: ctx
      type(phg_params), intent(in) :: prm
    end function

    type(c_ptr) function phg_last_error() bind(C, name='phg_last_error')
      import :: c_ptr
    end function

    ! ElmGMRs(y,ac,x,shp,shgl,iBC,BC,shpb,shglb,res,rmes,BDiag,iper,
    !         ilwork,lhsK,col,row,rerr)
    integer(c_int) function phg_elmgmrs(ctx, y, ac, res, rmes, BDiag, lhsK) &
        bind(C, name='phg_elmgmrs')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: y(*), ac(*)
      real(c_double), intent(out) :: res(*), rmes(*), BDiag(*), lhsK(*)
    end function

    ! SolGMRs(y,ac,yold,acold,x,iBC,BC,col,row,lhsk,res,BDiag,HBrg,eBrg,
    !         yBrg,Rcos,Rsin,iper,ilwork,shp,shgl,shpb,shglb,Dy,rerr)
    integer(c_int) function phg_solgmrs(ctx, y, ac, res, rmes, BDiag, lhsK, Dy, st) &
        bind(C, name='phg_solgmrs')
      import :: c_int, c_ptr, c_double, phg_stats
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: y(*), ac(*)
      real(c_double), intent(out) :: res(*), rmes(*), BDiag(*), Dy(*)
      type(c_ptr), value :: lhsK
      type(phg_stats), intent(out) :: st
    end function

    integer(c_int) function phg_state_upload(ctx, yold, acold) bind(C, name='phg_state_upload')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: yold(*), acold(*)
    end function

    integer(c_int) function phg_state_download(ctx, y, ac, Dy) bind(C, name='phg_state_download')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: y(*), ac(*), Dy(*)
    end function

    integer(c_int) function phg_state_download_old(ctx, yold, acold) &
        bind(C, name='phg_state_download_old')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: yold(*), acold(*)
    end function

    integer(c_int) function phg_state_upload_cur(ctx, y, ac) bind(C, name='phg_state_upload_cur')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: y(*), ac(*)
    end function

    ! itrPredict + itrBC
    integer(c_int) function phg_predict(ctx) bind(C, name='phg_predict')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    ! itrUpdate + itrBC on (yold, acold)
    integer(c_int) function phg_update(ctx) bind(C, name='phg_update')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function phg_newton_iteration(ctx, st) bind(C, name='phg_newton_iteration')
      import :: c_int, c_ptr, phg_stats
      type(c_ptr), value :: ctx
      type(phg_stats), intent(out) :: st
    end function

    ! SolMFG(y,ac,yold,acold,x,iBC,BC,res,BDiag,HBrg,eBrg,yBrg,Rcos,Rsin,
    !        iper,ilwork,shp,shgl,shpb,shglb,Dy,rerr): eGMRES is the
    ! /itrpar/ interval (in/out), fdi = 1 when iter = 1 and
    ! mod(istep, 20) = 0 (solmfg.f:163)
    integer(c_int) function phg_solmfg(ctx, y, ac, res, rmes, BDiag, Dy, eGMRES, fdi, st) &
        bind(C, name='phg_solmfg')
      import :: c_int, c_ptr, c_double, phg_stats
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: y(*), ac(*)
      real(c_double), intent(out) :: res(*), rmes(*), BDiag(*), Dy(*)
      real(c_double), intent(inout) :: eGMRES
      integer(c_int), value :: fdi
      type(phg_stats), intent(out) :: st
    end function

    integer(c_int) function phg_newton_iteration_mfg(ctx, eGMRES, fdi, st) &
        bind(C, name='phg_newton_iteration_mfg')
      import :: c_int, c_ptr, c_double, phg_stats
      type(c_ptr), value :: ctx
      real(c_double), intent(inout) :: eGMRES
      integer(c_int), value :: fdi
      type(phg_stats), intent(out) :: st
    end function
  end interface

contains

  ! Abort like the reference's error() when a call fails (error.f).
  subroutine phg_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    integer :: n
    if (rc == 0) return
    call c_f_pointer(phg_last_error(), msg, [1024])
    n = 0
    do while (n < 1024)
      if (msg(n + 1) == c_null_char) exit
      n = n + 1
    end do
    write(*, '(a,a,i4,a,1024a1)') where, ': libphgpu error ', rc, ': ', msg(1:n)
    error stop 1
  end subroutine

end module phgpu
