! iso_c_binding interface to the gfx950 engine (include/adpres_gpu.h).
!
! The derived types mirror the C structs field for field (bind(C) gives the
! same layout and padding); array arguments are passed as assumed-size
! arrays, i.e. by address, exactly as the C side expects Fortran
! column-major storage.
module adpres_gpu

  use iso_c_binding
  implicit none

  integer(c_int), parameter :: ADG_OK = 0, ADG_E_MAX_OUTER = 1, ADG_E_NODAL_UNSTABLE = 2, &
                               ADG_E_LU_PIVOT = 3, ADG_E_POWER_NONPOS = 4
  integer(c_int), parameter :: ADG_KERN_FDM = 0, ADG_KERN_PNM = 1, ADG_KERN_SANM = 2
  integer(c_int), parameter :: ADG_FORWARD = 0, ADG_ADJOINT = 1, ADG_FIXED_SOURCE = 2, ADG_TRANSIENT = 3
  integer(c_int), parameter :: ADG_ADJOINT_FROZEN = 4
  integer(c_int), parameter :: ADG_FP64 = 0, ADG_FP32_REFINE = 1
  integer(c_int), parameter :: ADG_E_ARG = 5, ADG_E_HIP = 6, ADG_E_NOT_SET = 7, ADG_E_XS = 8, &
                               ADG_E_DEVICE_SYNC = 9

  type, bind(C) :: adg_geom_t
    integer(c_int) :: nnod, ng, nmat, nxx, nyy, nzz
    type(c_ptr)    :: ix, iy, iz
    type(c_ptr)    :: xdel, ydel, zdel
    integer(c_int) :: bc(6)
    integer(c_int) :: precision
  end type

  type, bind(C) :: adg_iter_t
    integer(c_int) :: nout, nin, nac, nupd, kern
    real(c_double) :: serc, ferc
  end type

  type, bind(C) :: adg_trace_t
    integer(c_int) :: cap, n
    type(c_ptr)    :: k, ser, fer, ndmax, flags, loc
  end type

  ! feedback tables of XS_updt: tab(xs, kind) is C's tab[kind][xs]
  ! (kind 1 base .. 6 rod, xs 1 sigtr .. 5 sigs; c_null_ptr: feedback off)
  type, bind(C) :: adg_xs_tables_t
    integer(c_int) :: nmat, nb
    real(c_double) :: rbcon, rftem, rmtem, rcden, coreh, pos0, ssize
    type(c_ptr)    :: tab(5, 6)
    type(c_ptr)    :: fbmap
  end type

  ! branch tables of XStab_updt, flattened as include/adpres_gpu.h describes
  type, bind(C) :: adg_xtab_tables_t
    integer(c_int) :: nmat, nb
    type(c_ptr)    :: dims, points, xsec, rxsec
    real(c_double) :: coreh, pos0, ssize
    type(c_ptr)    :: fbmap
  end type

  interface

    integer(c_int) function adg_create(geom, ctx) bind(C, name='adg_create')
      import :: c_int, c_ptr, adg_geom_t
      type(adg_geom_t), intent(in) :: geom
      type(c_ptr), intent(out)     :: ctx
    end function

    subroutine adg_destroy(ctx) bind(C, name='adg_destroy')
      import :: c_ptr
      type(c_ptr), value :: ctx
    end subroutine

    integer(c_int) function adg_set_xs(ctx, D, sigr, sigs, nuf, sigf, chi, mat, dc, exsrc) &
        bind(C, name='adg_set_xs')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value         :: ctx
      real(c_double), intent(in) :: D(*), sigr(*), sigs(*), nuf(*), sigf(*), chi(*), dc(*), exsrc(*)
      integer(c_int), intent(in) :: mat(*)
    end function

    integer(c_int) function adg_set_state(ctx, f0, fs0, Ke, dn, first) bind(C, name='adg_set_state')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value    :: ctx, f0, fs0, dn
      real(c_double), value :: Ke
      integer(c_int), value :: first
    end function

    integer(c_int) function adg_get_state(ctx, f0, fs0, Ke, dn, df) bind(C, name='adg_get_state')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, f0, fs0, Ke, dn, df
    end function

    integer(c_int) function adg_outer(ctx, mode, iter, trace) bind(C, name='adg_outer')
      import :: c_int, c_ptr, adg_iter_t, adg_trace_t
      type(c_ptr), value              :: ctx
      integer(c_int), value           :: mode
      type(adg_iter_t), intent(in)    :: iter
      type(adg_trace_t), intent(inout):: trace
    end function

    integer(c_int) function adg_powdis(ctx, p) bind(C, name='adg_powdis')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value           :: ctx
      real(c_double), intent(out)  :: p(*)
    end function

    integer(c_int) function adg_last_timing(ctx, ms4) bind(C, name='adg_last_timing')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value          :: ctx
      real(c_double), intent(out) :: ms4(4)
    end function

    integer(c_int) function adg_set_profiling(ctx, on) bind(C, name='adg_set_profiling')
      import :: c_int, c_ptr
      type(c_ptr), value    :: ctx
      integer(c_int), value :: on
    end function

    type(c_ptr) function adg_last_error(ctx) bind(C, name='adg_last_error')
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function adg_set_transient(ctx, wtr) bind(C, name='adg_set_transient')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value          :: ctx
      real(c_double), intent(in)  :: wtr(*)
    end function

    integer(c_int) function adg_leakage(ctx, L) bind(C, name='adg_leakage')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value          :: ctx
      real(c_double), intent(out) :: L(*)
    end function

    integer(c_int) function adg_set_xs_tables(ctx, t, mat, chi, dc) bind(C, name='adg_set_xs_tables')
      import :: c_int, c_ptr, c_double, adg_xs_tables_t
      type(c_ptr), value                :: ctx
      type(adg_xs_tables_t), intent(in) :: t
      integer(c_int), intent(in)        :: mat(*)
      real(c_double), intent(in)        :: chi(*), dc(*)
    end function

    integer(c_int) function adg_set_xtab_tables(ctx, t, mat, chi) bind(C, name='adg_set_xtab_tables')
      import :: c_int, c_ptr, c_double, adg_xtab_tables_t
      type(c_ptr), value                  :: ctx
      type(adg_xtab_tables_t), intent(in) :: t
      integer(c_int), intent(in)          :: mat(*)
      real(c_double), intent(in)          :: chi(*)
    end function

    ! arrays of an inactive feedback may be c_null_ptr: passed by address
    integer(c_int) function adg_update_xs(ctx, bcon, ftem, mtem, cden, bpos) bind(C, name='adg_update_xs')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value    :: ctx, ftem, mtem, cden, bpos
      real(c_double), value :: bcon
    end function

    integer(c_int) function adg_get_xs(ctx, D, sigr, nuf, sigf, sigs, dc) bind(C, name='adg_get_xs')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, D, sigr, nuf, sigf, sigs, dc
    end function

    integer(c_int) function adg_power_maps(ctx, nx, ny, nz, xdiv, ydiv, zdiv, fasm, pmax, faxi, amax) &
        bind(C, name='adg_power_maps')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value          :: ctx
      integer(c_int), value       :: nx, ny, nz
      integer(c_int), intent(in)  :: xdiv(*), ydiv(*), zdiv(*)
      real(c_double), intent(out) :: fasm(*), faxi(*)
      integer(c_int), intent(out) :: pmax(2), amax
    end function

    integer(c_int) function adg_flux_map(ctx, nx, ny, xdiv, ydiv, norm, fasm, negf) bind(C, name='adg_flux_map')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value          :: ctx
      integer(c_int), value       :: nx, ny
      integer(c_int), intent(in)  :: xdiv(*), ydiv(*)
      real(c_double), value       :: norm
      real(c_double), intent(out) :: fasm(*)
      integer(c_int), intent(out) :: negf
    end function

  end interface

contains

  ! NUL-terminated C string -> Fortran string
  function c_string(p) result(s)
    type(c_ptr), intent(in) :: p
    character(len=200) :: s
    character(kind=c_char), pointer :: a(:)
    integer :: i
    s = ' '
    if (.not. c_associated(p)) return
    call c_f_pointer(p, a, [200])
    do i = 1, 200
      if (a(i) == c_null_char) exit
      s(i:i) = a(i)
    end do
  end function

end module adpres_gpu
