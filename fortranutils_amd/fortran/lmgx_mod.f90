! Fortran 2003 ISO_C_BINDING interface to liblmgx (the C ABI of include/lmgx.h).
! Types mirror lmgx_topology / lmgx_params field for field; every function returns 0 on success.
module lmgx_mod
  use iso_c_binding
  implicit none
  private
  public :: lmgx_topology, lmgx_params, lmgx_init, lmgx_free, lmgx_levels, lmgx_default_params, &
            lmgx_set_operator, lmgx_set_operator_cached, lmgx_counters, lmgx_solve, &
            lmgx_uvw_adv, lmgx_last_error_str, lmgx_shared, lmgx_force_upload, &
            LMGX_SMOOTH_EXACT, LMGX_SMOOTH_LINE, LMGX_SMOOTH_SPORLX2, LMGX_PREC_FP64, &
            LMGX_PREC_MIXED

  ! the one handle of a LESter run, created on first use by lmgx_attach (lmg_solve_gpu.F) and
  ! shared by the drop-ins lmg_solve and uvw_adv
  type(c_ptr), save :: lmgx_shared = c_null_ptr
  ! set by lmgx_coeff_changed (lmg_solve_gpu.F): the next lmg_solve uploads the coefficients
  ! even when their arrays and sampled fingerprint are those of the last upload
  integer(c_int), save :: lmgx_force_upload = 0

  integer(c_int), parameter :: LMGX_SMOOTH_EXACT = 0, LMGX_SMOOTH_LINE = 1, LMGX_SMOOTH_SPORLX2 = 2
  integer(c_int), parameter :: LMGX_PREC_FP64 = 0, LMGX_PREC_MIXED = 1

  type, bind(c) :: lmgx_topology
    integer(c_int) :: nblk, nlev
    type(c_ptr) :: nijkp1, bc, bl2bl, bl2side
  end type

  type, bind(c) :: lmgx_params
    integer(c_int) :: cycle, numpre, numpost, numcits, max_cycles
    real(c_double) :: omega, omega_c, tolf, rtol
    integer(c_int) :: smoother, precision, check_every, sweeps_per_iter, coarse_op, coarse_solve
    integer(c_int) :: fine_post
  end type

  interface
    integer(c_int) function lmgx_init(topo, h) bind(c, name='lmgx_init')
      import :: c_int, c_ptr, lmgx_topology
      type(lmgx_topology), intent(in) :: topo
      type(c_ptr), intent(out) :: h
    end function

    integer(c_int) function lmgx_free(h) bind(c, name='lmgx_free')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function

    integer(c_int) function lmgx_levels(h, nlev, ntot, n1, dims, off) bind(c, name='lmgx_levels')
      import :: c_int, c_ptr, c_long
      type(c_ptr), value :: h
      integer(c_int), intent(out) :: nlev
      integer(c_long), intent(out) :: ntot, n1
      integer(c_int), intent(out) :: dims(*)
      integer(c_long), intent(out) :: off(*)
    end function

    subroutine lmgx_default_params(p) bind(c, name='lmgx_default_params')
      import :: lmgx_params
      type(lmgx_params), intent(out) :: p
    end subroutine

    integer(c_int) function lmgx_set_operator(h, ap, aw, ae, as, an, al, ar) &
        bind(c, name='lmgx_set_operator')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: ap(*), aw(*), ae(*), as(*), an(*), al(*), ar(*)
    end function

    integer(c_int) function lmgx_set_operator_cached(h, ap, aw, ae, as, an, al, ar, force, &
        uploaded) bind(c, name='lmgx_set_operator_cached')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: ap(*), aw(*), ae(*), as(*), an(*), al(*), ar(*)
      integer(c_int), value :: force
      integer(c_int), intent(out) :: uploaded
    end function

    integer(c_int) function lmgx_counters(h, cnt) bind(c, name='lmgx_counters')
      import :: c_int, c_ptr, c_long
      type(c_ptr), value :: h
      integer(c_long), intent(out) :: cnt(4)
    end function

    integer(c_int) function lmgx_solve(h, f, b, residu, p, hist, ncyc) bind(c, name='lmgx_solve')
      import :: c_int, c_ptr, c_double, lmgx_params
      type(c_ptr), value :: h
      real(c_double), intent(inout) :: f(*), b(*), residu(*)
      type(lmgx_params), intent(in) :: p
      real(c_double), intent(out) :: hist(*)
      integer(c_int), intent(out) :: ncyc
    end function

    integer(c_int) function lmgx_uvw_adv(h, ivar, diu, rp, dkw, u, u_int, rho, p, dtime) &
        bind(c, name='lmgx_uvw_adv')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: ivar
      real(c_double), intent(in) :: diu(*), rp(*), dkw(*), u_int(*), rho(*), p(*)
      real(c_double), intent(inout) :: u(*)
      real(c_double), value :: dtime
    end function

    type(c_ptr) function lmgx_last_error() bind(c, name='lmgx_last_error')
      import :: c_ptr
    end function

    integer(c_size_t) function c_strlen(s) bind(c, name='strlen')
      import :: c_size_t, c_ptr
      type(c_ptr), value :: s
    end function
  end interface

contains

  ! the library's last error message as a Fortran string
  function lmgx_last_error_str() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: cs(:)
    integer :: n, i
    p = lmgx_last_error()
    n = int(c_strlen(p))
    call c_f_pointer(p, cs, [n])
    allocate(character(len=n) :: msg)
    do i = 1, n
      msg(i:i) = cs(i)
    end do
  end function

end module lmgx_mod
