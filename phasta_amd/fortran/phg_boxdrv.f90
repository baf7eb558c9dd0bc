!-----------------------------------------------------------------------
! phg_boxdrv: Fortran host that drives libphgpu.so through the module
! phgpu, in the shape of the reference's time loop (itrdrv.f:398-652):
!
!   do istep:  itrPredict + itrBC
!              do iter = 1, nitr:  SolGMRs -> itrCorrect + itrBC
!              itrUpdate + itrBC
!
! Two ways of calling the library are exercised:
!   mode 'host'     -- the drop-in: the host keeps y, ac, yold, acold in its
!                      own arrays and replaces only the SolGMRs call
!                      (itrdrv.f:477) by phg_solgmrs; predictor, corrector
!                      and BCs stay host code written here.
!   mode 'resident' -- the state lives in HBM; the host only sequences
!                      phg_predict / phg_newton_iteration / phg_update.
!
! Usage: phg_boxdrv N mach kspace steps newton mode outfile
! Writes yold(nshg,5) after the last step to `outfile` (raw float64,
! column-major) and one line per Newton iteration to stdout.
!-----------------------------------------------------------------------
module phmesh
  use iso_c_binding
  implicit none
  interface
    integer(c_int) function phm_box_counts(N, nshg, nel, nelb) bind(C, name='phm_box_counts')
      import :: c_int, c_long_long
      integer(c_int), value :: N
      integer(c_long_long), intent(out) :: nshg, nel, nelb
    end function
    integer(c_int) function phm_box_mesh(N, jitter, grade, seed, x, ien, ienb, iBC, iper) &
        bind(C, name='phm_box_mesh')
      import :: c_int, c_double, c_long_long
      integer(c_int), value :: N
      real(c_double), value :: jitter, grade
      integer(c_long_long), value :: seed
      real(c_double), intent(out) :: x(*)
      integer(c_int), intent(out) :: ien(*), ienb(*), iBC(*), iper(*)
    end function
    integer(c_int) function phm_csr_count(nshg, nel, ien, colm, nnz) bind(C, name='phm_csr_count')
      import :: c_int, c_long_long
      integer(c_long_long), value :: nshg, nel
      integer(c_int), intent(in) :: ien(*)
      integer(c_int), intent(out) :: colm(*)
      integer(c_long_long), intent(out) :: nnz
    end function
    integer(c_int) function phm_csr_fill(nshg, nel, ien, colm, rowp) bind(C, name='phm_csr_fill')
      import :: c_int, c_long_long
      integer(c_long_long), value :: nshg, nel
      integer(c_int), intent(in) :: ien(*), colm(*)
      integer(c_int), intent(out) :: rowp(*)
    end function
    integer(c_int) function phm_box_state(nshg, x, mach, seed, Rgas, gamma, y) &
        bind(C, name='phm_box_state')
      import :: c_int, c_double, c_long_long
      integer(c_long_long), value :: nshg
      real(c_double), intent(in) :: x(*)
      real(c_double), value :: mach
      integer(c_long_long), value :: seed
      real(c_double), value :: Rgas, gamma
      real(c_double), intent(out) :: y(*)
    end function
  end interface
end module phmesh

program phg_boxdrv
  use iso_c_binding
  use phgpu
  use phmesh
  implicit none

  integer(c_long_long), parameter :: mesh_seed = 20261202_c_long_long
  integer(c_long_long), parameter :: field_seed = 7_c_long_long
  integer, parameter :: ndofBC = 10
  real(c_double), parameter :: Rh = 8.31441d0, jitter = 0.1d0, grade = 1.0d0
  real(c_double), parameter :: T0 = 300.0d0, cfl = 1.0d0

  integer :: N, kspace, nsteps, nitr, istep, iter, i
  real(c_double) :: mach, Rgas, c0, dt
  character(len=256) :: arg, mode, outfile
  integer(c_long_long) :: nshg, nel, nelb, nnz
  real(c_double), allocatable, target :: x(:,:), BC(:,:)
  integer(c_int), allocatable, target :: ien(:,:), ienb(:,:), iBC(:), iper(:), colm(:), rowp(:)
  real(c_double), allocatable :: y(:,:), ac(:,:), yold(:,:), acold(:,:)
  real(c_double), allocatable :: res(:,:), rmes(:,:), BDiag(:,:,:), Dy(:,:)
  type(phg_mesh) :: mesh
  type(phg_params) :: prm
  type(phg_stats) :: st
  type(c_ptr) :: ctx

  if (command_argument_count() /= 7) then
    write(*, '(a)') 'usage: phg_boxdrv N mach kspace steps newton host|resident outfile'
    stop 2
  end if
  call get_command_argument(1, arg); read(arg, *) N
  call get_command_argument(2, arg); read(arg, *) mach
  call get_command_argument(3, arg); read(arg, *) kspace
  call get_command_argument(4, arg); read(arg, *) nsteps
  call get_command_argument(5, arg); read(arg, *) nitr
  call get_command_argument(6, mode)
  call get_command_argument(7, outfile)
  if (mode /= 'host' .and. mode /= 'resident') then
    write(*, '(2a)') 'unknown mode ', trim(mode)
    stop 2
  end if

  ! ---- mesh and initial state (stand-in for readnblk/genblk/genadj) ----
  if (phm_box_counts(N, nshg, nel, nelb) /= 0) stop 'bad N'
  allocate(x(nshg,3), ien(nel,4), ienb(nelb,4), iBC(nshg), iper(nshg))
  allocate(colm(nshg+1), BC(nshg,ndofBC))
  if (phm_box_mesh(N, jitter, grade, mesh_seed, x, ien, ienb, iBC, iper) /= 0) stop 'mesh'
  if (phm_csr_count(nshg, nel, ien, colm, nnz) /= 0) stop 'csr count'
  allocate(rowp(nnz))
  if (phm_csr_fill(nshg, nel, ien, colm, rowp) /= 0) stop 'csr fill'
  BC = 0.0d0

  ! gas constant of the N2/O2 mix (input.f:165-182), Dtgl = 1/dt
  Rgas = 1.0d0 / (0.79d0 / (Rh / 2.8d-2) + 0.21d0 / (Rh / 3.2d-2))
  c0 = sqrt(1.4d0 * Rgas * T0)
  dt = cfl * (1.0d0 / N) / (c0 * (1.0d0 + mach))

  prm%Rgas = Rgas;  prm%gamma = 1.4d0;  prm%gamma1 = 0.4d0
  prm%pr = 0.72d0;  prm%mu = 1.8d-5
  prm%dtsfct = 1.0d0;  prm%taucfct = 1.0d0;  prm%temper = 1.0d0
  prm%almi = 1.0d0;  prm%alfi = 1.0d0;  prm%gami = 1.0d0
  prm%Dtgl = 1.0d0 / dt
  prm%etol = 1.0d-3;  prm%epsM = sqrt(epsilon(1.0d0))
  prm%Kspace = kspace;  prm%nGMRES = 1;  prm%minIters = 0
  prm%lhs = 1;  prm%iprec = 1;  prm%ipred = 1;  prm%idiff = 0;  prm%iter = 0;  prm%idc = 0;  prm%itau = 0

  allocate(y(nshg,5), ac(nshg,5), yold(nshg,5), acold(nshg,5))
  allocate(res(nshg,5), rmes(nshg,5), BDiag(nshg,5,5), Dy(nshg,5))
  if (phm_box_state(nshg, x, mach, field_seed, Rgas, prm%gamma, yold) /= 0) stop 'state'
  acold = 0.0d0

  mesh%nshg = nshg;  mesh%nel = nel;  mesh%nelb = nelb;  mesh%nnz_tot = nnz
  mesh%ndofBC = ndofBC;  mesh%reserved = 0
  mesh%x = c_loc(x);  mesh%ien = c_loc(ien);  mesh%ienb = c_loc(ienb)
  mesh%iBC = c_loc(iBC);  mesh%BC = c_loc(BC);  mesh%iper = c_loc(iper)
  mesh%colm = c_loc(colm);  mesh%rowp = c_loc(rowp)
  call phg_check(phg_init(mesh, prm, PHG_FP64, ctx), 'phg_init')

  if (mode == 'resident') then
    call phg_check(phg_state_upload(ctx, yold, acold), 'phg_state_upload')
  end if

  do istep = 1, nsteps
    if (mode == 'host') then
      call predict_bc()
    else
      call phg_check(phg_predict(ctx), 'phg_predict')
    end if
    do iter = 1, nitr
      if (mode == 'host') then
        ! /timdat/ iter (itrdrv.f:409,434), read by rstatCheck inside SolGMRs
        prm%iter = iter
        call phg_check(phg_set_params(ctx, prm), 'phg_set_params')
        ! itrdrv.f:477 call SolGMRs(y, ac, yold, acold, x, iBC, BC, col, row,
        !              lhsk, res, BDiag, ..., iper, ilwork, shp, ..., Dy, rerr)
        call phg_check(phg_solgmrs(ctx, y, ac, res, rmes, BDiag, c_null_ptr, Dy, st), &
                       'SolGMRs')
        call correct_bc()
      else
        call phg_check(phg_newton_iteration(ctx, st), 'phg_newton_iteration')
      end if
      write(*, '(a,i4,a,i3,a,i4,a,es22.14,a,es22.14)') 'step', istep, ' iter', iter, &
          ' iKs', st%iKs, ' res', st%totres(1), ' unorm0', st%unorm0
    end do
    if (mode == 'host') then
      call update_bc()
    else
      call phg_check(phg_update(ctx), 'phg_update')
    end if
  end do

  if (mode == 'resident') then
    call phg_check(phg_state_download_old(ctx, yold, acold), 'phg_state_download_old')
  end if
  call phg_check(phg_finalize(ctx), 'phg_finalize')

  open(unit=21, file=trim(outfile), access='stream', form='unformatted', status='replace')
  write(21) yold
  close(21)

contains

  ! itrPredict with ipred = 1 (itrPC.f:67-73), then itrBC
  subroutine predict_bc()
    y = yold
    ac = acold * (1.0d0 - prm%almi / prm%gami)
    call bc_y(y, ac)
  end subroutine

  ! itrCorrect (itrPC.f:128-148): Dy is in equation order (p,u,v,w,T)
  subroutine correct_bc()
    real(c_double) :: fct1, fct2
    y(:,1:3) = y(:,1:3) - Dy(:,2:4)
    y(:,4) = y(:,4) - Dy(:,1)
    y(:,5) = y(:,5) - Dy(:,5)
    fct1 = 1.0d0 - prm%almi / prm%gami
    fct2 = prm%almi * prm%Dtgl / prm%gami / prm%alfi
    ac = acold * fct1 + (y - yold) * fct2
    call bc_y(y, ac)
  end subroutine

  ! itrUpdate (itrPC.f:204-207), then itrBC on (yold, acold)
  subroutine update_bc()
    acold = acold + (ac - acold) * (1.0d0 / prm%almi)
    yold = yold + (y - yold) * (1.0d0 / prm%alfi)
    call bc_y(yold, acold)
  end subroutine

  ! itrBC for the codes of the box: no-slip walls (velocity code 7, BC
  ! slots 3-5) and periodic slaves copied from their final masters.
  subroutine bc_y(yy, aa)
    real(c_double), intent(inout) :: yy(nshg,5), aa(nshg,5)
    integer :: f, k
    do k = 1, int(nshg)
      if (iand(ishft(iBC(k), -3), 7) == 7) yy(k,1:3) = BC(k,3:5)
      if (iand(iBC(k), 2) /= 0) yy(k,5) = BC(k,2)
      if (iand(iBC(k), 4) /= 0) yy(k,4) = BC(k,1)
    end do
    do f = 1, 5
      do k = 1, int(nshg)
        yy(k,f) = yy(iper(k),f)
        aa(k,f) = aa(iper(k),f)
      end do
    end do
  end subroutine

end program phg_boxdrv
