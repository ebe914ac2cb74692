! bertha_amd_mod.f90 -- ISO_C_BINDING interfaces to include/bertha_amd.h so a
! Fortran SCF driver (the reference's HFSCF, bertha-scf.f:5097) can call the
! engine in place of CALL EQSAVE / CALL CLMFAST without changing its own data:
! COMPLEX*16 (MDM,MDM) arrays pass straight through with LD = MDM.
module bertha_amd
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: BDF_OK = 0
  integer(c_int), parameter :: BDF_SHAPE_NOSYMM = 0, BDF_SHAPE_LINEAR = 1, BDF_SHAPE_ATOMIC = 2
  integer(c_int), parameter :: BDF_HMLT_DHFR = 1, BDF_HMLT_DHFB = 2
  integer(c_int), parameter :: BDF_PREC_FP64 = 0, BDF_PREC_MIXED = 1
  integer(c_int), parameter :: BDF_SCHWARZ_OFF = 0, BDF_SCHWARZ_REFERENCE = 1, BDF_SCHWARZ_RIGOROUS = 2

  type, bind(c) :: bdf_system
    integer(c_int) :: ncnt
    type(c_ptr) :: xyz, lmax, nfnc, exps
    integer(c_int) :: lstride, mbs, shape, hmlt
  end type bdf_system

  type, bind(c) :: bdf_options
    integer(c_int) :: intsym, racah1, precision
    real(c_double) :: split_threshold
    integer(c_int) :: device, schwarz
    real(c_double) :: sens
    integer(c_int) :: gaunt, workgroups
  end type bdf_options

  type, bind(c) :: bdf_launch
    integer(c_int) :: graph, waves, cap, per_cu, breit_per_cu, breit_group, breit_kernel
  end type bdf_launch

  interface
    integer(c_int) function bdf_create(sys, opt, eng, ndim, nskp) bind(c, name='bdf_create')
      import :: c_int, c_ptr, bdf_system, bdf_options
      type(bdf_system), intent(in) :: sys
      type(bdf_options), intent(in) :: opt
      type(c_ptr), intent(out) :: eng
      integer(c_int), intent(out) :: ndim, nskp
    end function

    integer(c_int) function bdf_coulomb(eng, ilev, dent, deno, ld, gdir, gxch, qdir, qxch) &
        bind(c, name='bdf_coulomb')
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: eng
      integer(c_int), value :: ilev, ld
      complex(c_double_complex), intent(in) :: dent(ld, *)
      type(c_ptr), value :: deno, qdir, qxch
      complex(c_double_complex), intent(out) :: gdir(ld, *), gxch(ld, *)
    end function

    integer(c_int) function bdf_coulomb_device(eng, ilev, d_dent, d_gdir, d_gxch) &
        bind(c, name='bdf_coulomb_device')
      import :: c_int, c_ptr
      type(c_ptr), value :: eng, d_dent, d_gdir, d_gxch
      integer(c_int), value :: ilev
    end function

    integer(c_int) function bdf_coulomb_device_static(eng, ilev, d_dent, d_gdir, d_gxch) &
        bind(c, name='bdf_coulomb_device_static')
      import :: c_int, c_ptr
      type(c_ptr), value :: eng, d_dent, d_gdir, d_gxch
      integer(c_int), value :: ilev
    end function

    ! in place of DO IZ=1,NCNT: CALL COULOMB1(IZ) (bertha-scf.f:5415-5421)
    integer(c_int) function bdf_coulomb1(eng, exact_beta, dent, deno, ld, gdir, gxch, qdir, qxch) &
        bind(c, name='bdf_coulomb1')
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: eng
      integer(c_int), value :: exact_beta, ld
      complex(c_double_complex), intent(in) :: dent(ld, *)
      type(c_ptr), value :: deno, qdir, qxch
      complex(c_double_complex), intent(inout) :: gdir(ld, *), gxch(ld, *)
    end function

    integer(c_int) function bdf_coulomb1_device(eng, exact_beta, d_dent, d_gdir, d_gxch) &
        bind(c, name='bdf_coulomb1_device')
      import :: c_int, c_ptr
      type(c_ptr), value :: eng, d_dent, d_gdir, d_gxch
      integer(c_int), value :: exact_beta
    end function

    ! in place of DO IZ=1,NCNT: CALL BREIT1(IZ) (bertha-scf.f:5447-5453)
    integer(c_int) function bdf_breit1(eng, exact_beta, dent, deno, ld, bdir, bxch, wdir, wxch) &
        bind(c, name='bdf_breit1')
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: eng
      integer(c_int), value :: exact_beta, ld
      complex(c_double_complex), intent(in) :: dent(ld, *)
      type(c_ptr), value :: deno, wdir, wxch
      complex(c_double_complex), intent(inout) :: bdir(ld, *), bxch(ld, *)
    end function

    integer(c_int) function bdf_breit1_device(eng, exact_beta, d_dent, d_bdir, d_bxch) &
        bind(c, name='bdf_breit1_device')
      import :: c_int, c_ptr
      type(c_ptr), value :: eng, d_dent, d_bdir, d_bxch
      integer(c_int), value :: exact_beta
    end function

    integer(c_int) function bdf_racah_info(eng, breit, terms, items, uncovered) &
        bind(c, name='bdf_racah_info')
      import :: c_int, c_ptr, c_double, c_long
      type(c_ptr), value :: eng
      integer(c_int), value :: breit
      real(c_double), intent(out) :: terms
      integer(c_long), intent(out) :: items, uncovered
    end function

    integer(c_int) function bdf_eri_batch(eng, icnt, kidx, midx, ibas, jbas, itn, ilev, rr, maxcd) &
        bind(c, name='bdf_eri_batch')
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: eng
      integer(c_int), intent(in) :: icnt(4), kidx(4), midx(4), itn(2)
      integer(c_int), value :: ibas, jbas, ilev, maxcd
      complex(c_double_complex), intent(out) :: rr(maxcd, 16)
    end function

    integer(c_int) function bdf_etable(eng, itt, ica, icb, ka, kb, ma, mb, out, cap, n) &
        bind(c, name='bdf_etable')
      import :: c_int, c_ptr, c_double, c_long
      type(c_ptr), value :: eng
      integer(c_int), value :: itt, ica, icb, ka, kb, ma, mb
      real(c_double), intent(out) :: out(*)
      integer(c_long), value :: cap
      integer(c_long), intent(out) :: n
    end function

    integer(c_int) function bdf_breit(eng, ilev, dent, deno, ld, bdir, bxch, wdir, wxch) &
        bind(c, name='bdf_breit')
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: eng
      integer(c_int), value :: ilev, ld
      complex(c_double_complex), intent(in) :: dent(ld, *)
      type(c_ptr), value :: deno, wdir, wxch
      complex(c_double_complex), intent(out) :: bdir(ld, *), bxch(ld, *)
    end function

    integer(c_int) function bdf_breit_device(eng, ilev, d_dent, d_bdir, d_bxch) &
        bind(c, name='bdf_breit_device')
      import :: c_int, c_ptr
      type(c_ptr), value :: eng, d_dent, d_bdir, d_bxch
      integer(c_int), value :: ilev
    end function

    integer(c_int) function bdf_bii_batch(eng, icnt, kidx, midx, ibas, jbas, rr, maxcd) &
        bind(c, name='bdf_bii_batch')
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: eng
      integer(c_int), intent(in) :: icnt(4), kidx(4), midx(4)
      integer(c_int), value :: ibas, jbas, maxcd
      complex(c_double_complex), intent(out) :: rr(maxcd, 16)
    end function

    integer(c_int) function bdf_breit_plan_info(eng, ilev, flops, quartets, items) &
        bind(c, name='bdf_breit_plan_info')
      import :: c_int, c_ptr, c_double, c_long
      type(c_ptr), value :: eng
      integer(c_int), value :: ilev
      real(c_double), intent(out) :: flops, quartets
      integer(c_long), intent(out) :: items
    end function

    integer(c_int) function bdf_one_electron(eng, znuc, xi, cv, ovlp, hnuc, hkin, ld) &
        bind(c, name='bdf_one_electron')
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: eng
      real(c_double), intent(in) :: znuc(*), xi(*)
      real(c_double), value :: cv
      integer(c_int), value :: ld
      complex(c_double_complex), intent(out) :: ovlp(ld, *), hnuc(ld, *), hkin(ld, *)
    end function

    integer(c_int) function bdf_one_electron_device(eng, znuc, xi, cv, d_ovlp, d_hnuc, d_hkin) &
        bind(c, name='bdf_one_electron_device')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: eng, d_ovlp, d_hnuc, d_hkin
      real(c_double), intent(in) :: znuc(*), xi(*)
      real(c_double), value :: cv
    end function

    integer(c_int) function bdf_schwarz_tables(eng, gdsc, qbnd, ld) &
        bind(c, name='bdf_schwarz_tables')
      import :: c_int, c_ptr
      type(c_ptr), value :: eng, gdsc, qbnd
      integer(c_int), value :: ld
    end function

    integer(c_int) function bdf_memory_info(eng, device_bytes) bind(c, name='bdf_memory_info')
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: eng
      integer(c_size_t), intent(out) :: device_bytes
    end function

    type(c_ptr) function bdf_stream(eng) bind(c, name='bdf_stream')
      import :: c_ptr
      type(c_ptr), value :: eng
    end function

    integer(c_int) function bdf_last_stats(eng, quartets, kernel_ms) bind(c, name='bdf_last_stats')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: eng
      real(c_double), intent(out) :: quartets, kernel_ms
    end function

    integer(c_int) function bdf_plan_info(eng, ilev, flops, quartets, items) &
        bind(c, name='bdf_plan_info')
      import :: c_int, c_ptr, c_double, c_long
      type(c_ptr), value :: eng
      integer(c_int), value :: ilev
      real(c_double), intent(out) :: flops, quartets
      integer(c_long), intent(out) :: items
    end function

    subroutine bdf_destroy(eng) bind(c, name='bdf_destroy')
      import :: c_ptr
      type(c_ptr), value :: eng
    end subroutine

    integer(c_int) function bdf_set_launch(eng, l) bind(c, name='bdf_set_launch')
      import :: c_int, c_ptr, bdf_launch
      type(c_ptr), value :: eng
      type(bdf_launch), intent(in) :: l
    end function

    integer(c_int) function bdf_breit_launch_info(eng, ilev, kernel, workgroups) &
        bind(c, name='bdf_breit_launch_info')
      import :: c_int, c_ptr
      type(c_ptr), value :: eng
      integer(c_int), value :: ilev
      integer(c_int), intent(out) :: kernel, workgroups
    end function

    integer(c_int) function bdf_launch_info(eng, ilev, waves, cap, workgroups, scratch_bytes) &
        bind(c, name='bdf_launch_info')
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: eng
      integer(c_int), value :: ilev
      integer(c_int), intent(out) :: waves, cap, workgroups
      integer(c_size_t), intent(out) :: scratch_bytes
    end function

    type(c_ptr) function bdf_last_error() bind(c, name='bdf_last_error')
      import :: c_ptr
    end function
  end interface
end module bertha_amd
