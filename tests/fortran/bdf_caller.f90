! bdf_caller.f90 -- a Fortran driver of the engine through the ISO_C_BINDING
! module include/bertha_amd_mod.f90, called the way HFSCF calls CLMFAST
! (bertha-scf.f:5401): COMPLEX*16 arrays with a leading dimension MDM larger
! than NDIM pass straight through.  Built and run by tests/test_fortran.py.
!
!   bdf_caller abi         struct sizes and error codes (no GPU needed)
!   bdf_caller ne DIR      Ne 12s8p (alpha_k = 0.08 * 2.6**k): reads the
!                          density DIR/dens.bin (NDIM x NDIM, column-major,
!                          stream), writes GDIR and GXCH of a stage-3 build
!                          to DIR/gx.bin
program bdf_caller
  use iso_c_binding
  use bertha_amd
  implicit none
  integer, parameter :: mdm_pad = 3
  character(len=16) :: mode
  character(len=512) :: dir
  type(bdf_system) :: sys
  type(bdf_options) :: opt
  type(c_ptr) :: eng
  integer(c_int) :: rc, ndim, nskp, waves, cap, wgs
  integer(c_size_t) :: scratch
  real(c_double), target :: xyz(3, 1), exps(32, 4, 1)
  integer(c_int), target :: lmax(1), nfnc(4, 1)
  complex(c_double_complex), allocatable :: d(:, :), g(:, :), x(:, :), dn(:, :)
  integer :: k, ld, u

  call get_command_argument(1, mode)
  xyz = 0.0d0
  lmax(1) = 1
  nfnc = 0
  nfnc(1, 1) = 12
  nfnc(2, 1) = 8
  exps = 0.0d0
  do k = 1, 12
    exps(k, 1, 1) = 0.08d0 * 2.6d0**(k - 1)
  end do
  do k = 1, 8
    exps(k, 2, 1) = 0.08d0 * 2.6d0**(k - 1)
  end do
  sys%ncnt = 1
  sys%xyz = c_loc(xyz)
  sys%lmax = c_loc(lmax)
  sys%nfnc = c_loc(nfnc)
  sys%exps = c_loc(exps)
  sys%lstride = 4
  sys%mbs = 32
  sys%shape = BDF_SHAPE_ATOMIC
  sys%hmlt = BDF_HMLT_DHFR
  opt = bdf_options(1, 0, BDF_PREC_FP64, 0.0d0, 0, BDF_SCHWARZ_OFF, 0.0d0, 0, 0)

  if (trim(mode) == 'abi') then
    print '(a,2i6)', 'sizes', c_sizeof(sys), c_sizeof(opt)
    allocate(dn(2, 2))
    dn = (0.0d0, 0.0d0)
    rc = bdf_coulomb(c_null_ptr, 3, dn, c_null_ptr, 2, dn, dn, c_null_ptr, c_null_ptr)
    print '(a,i4)', 'coulomb_null_engine', rc
    rc = bdf_launch_info(c_null_ptr, 3, waves, cap, wgs, scratch)
    print '(a,i4)', 'launch_info_null_engine', rc
    sys%hmlt = 7
    rc = bdf_create(sys, opt, eng, ndim, nskp)
    print '(a,i4)', 'create_bad_hamiltonian', rc
    sys%hmlt = BDF_HMLT_DHFR
    opt%precision = 5
    rc = bdf_create(sys, opt, eng, ndim, nskp)
    print '(a,i4)', 'create_bad_precision', rc
    stop
  end if

  call get_command_argument(2, dir)
  rc = bdf_create(sys, opt, eng, ndim, nskp)
  if (rc /= BDF_OK) then
    print '(a,i4)', 'create failed', rc
    stop 1
  end if
  ld = ndim + mdm_pad
  allocate(d(ld, ndim), g(ld, ndim), x(ld, ndim))
  d = (0.0d0, 0.0d0)
  open(newunit=u, file=trim(dir)//'/dens.bin', access='stream', form='unformatted', status='old')
  read(u) d(1:ndim, 1:ndim)
  close(u)
  ! the padding rows must come back untouched
  d(ndim + 1:ld, :) = (7.0d0, -7.0d0)
  g = (7.0d0, -7.0d0)
  x = (7.0d0, -7.0d0)
  rc = bdf_coulomb(eng, 3, d, c_null_ptr, ld, g, x, c_null_ptr, c_null_ptr)
  if (rc /= BDF_OK) then
    print '(a,i4)', 'coulomb failed', rc
    stop 1
  end if
  if (any(g(ndim + 1:ld, :) /= (7.0d0, -7.0d0)) .or. any(x(ndim + 1:ld, :) /= (7.0d0, -7.0d0))) then
    print '(a)', 'padding rows overwritten'
    stop 1
  end if
  open(newunit=u, file=trim(dir)//'/gx.bin', access='stream', form='unformatted', status='replace')
  write(u) g(1:ndim, 1:ndim), x(1:ndim, 1:ndim)
  close(u)
  call bdf_destroy(eng)
  print '(a,2i6)', 'ne ok', ndim, nskp
end program bdf_caller
