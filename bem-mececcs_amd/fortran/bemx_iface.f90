! bemx_iface.f90 -- iso_c_binding interfaces to libbemx.so (include/bemx.h)
! and the device context shared by the drop-in subroutines.
!
! The reference kept its device state in module globals of
! kernels/shared.cu:6-31; here one context handle per process plays that
! role.  Errors come back as status codes; bemx_check prints the library's
! message and stops, as the reference's Fortran stops on its own errors
! (e.g. Nonsingd.for:89-95).
module bemx_iface
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: BEMX_MODE_FP64 = 0, BEMX_MODE_MIXED = 1

  interface
    integer(c_int) function bemx_create(device, ctx) bind(C, name='bemx_create')
      import :: c_int, c_ptr
      integer(c_int), value :: device
      type(c_ptr) :: ctx
    end function
    integer(c_int) function bemx_destroy(ctx) bind(C, name='bemx_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    type(c_ptr) function bemx_last_error(ctx) bind(C, name='bemx_last_error')
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function bemx_gauleg(npg, gi, ome) bind(C, name='bemx_gauleg')
      import :: c_int, c_double
      integer(c_int), value :: npg
      real(c_double) :: gi(*), ome(*)
    end function
    integer(c_int) function bemx_send_shared_data(ctx, cx, cy, cz, cxm, cym, czm, etas, gi, &
        ome, cone, np, npg, n, nbe) bind(C, name='bemx_send_shared_data')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: cx(*), cy(*), cz(*), cxm(*), cym(*), czm(*), etas(*), gi(*), ome(*)
      integer(c_int) :: cone(*)
      integer(c_int), value :: np, npg, n, nbe
    end function
    integer(c_int) function bemx_deallocate_shared_data(ctx) &
        bind(C, name='bemx_deallocate_shared_data')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function bemx_ghmatece(ctx, ge, rnu, hestdiag, gestdiag, delta, pi, c1, c2, &
        c3, c4) bind(C, name='bemx_ghmatece')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: ge, rnu
      real(c_double) :: hestdiag(*), gestdiag(*), delta(*), pi, c1, c2, c3, c4
    end function
    integer(c_int) function bemx_send_dyn_data(ctx, zdfi, kode) &
        bind(C, name='bemx_send_dyn_data')
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex) :: zdfi(*)
      integer(c_int) :: kode(*)
    end function
    integer(c_int) function bemx_ghmatecd(ctx, zge, zcs, zcp, c1, c2, c3, c4, fr, hestdiag, &
        gestdiag, zfi, zh) bind(C, name='bemx_ghmatecd')
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex) :: zge, zcs, zcp
      real(c_double), value :: c1, c2, c3, c4, fr
      real(c_double) :: hestdiag(*), gestdiag(*)
      complex(c_double_complex) :: zfi(*), zh(*)
    end function
    integer(c_int) function bemx_linsolve(ctx, nn, zh, zfi, mode, info) &
        bind(C, name='bemx_linsolve')
      import :: c_int, c_ptr, c_double_complex, c_long_long
      type(c_ptr), value :: ctx
      integer(c_int), value :: nn, mode
      complex(c_double_complex) :: zh(*), zfi(*)
      integer(c_long_long) :: info
    end function
    integer(c_int) function bemx_interec(ctx, zfi, zdfi, kode, l, cxi, cyi, czi, zge, zcs, zcp, &
        fr, rho, zdsol, zssol) bind(C, name='bemx_interec')
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex) :: zfi(*), zdfi(*)
      integer(c_int) :: kode(*)
      integer(c_int), value :: l
      real(c_double) :: cxi(*), cyi(*), czi(*)
      complex(c_double_complex) :: zge, zcs, zcp
      real(c_double), value :: fr, rho
      complex(c_double_complex) :: zdsol(*), zssol(*)
    end function
  end interface

  type(c_ptr), save :: bemx_ctx = c_null_ptr
  logical, save :: bemx_mesh_sent = .false.
  integer(c_int), save :: bemx_solver_mode = BEMX_MODE_FP64

contains

  subroutine bemx_check(code, what)
    integer(c_int), intent(in) :: code
    character(len=*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    type(c_ptr) :: p
    integer :: k
    if (code == 0) return
    write(0, '(A,A,A,I0)') 'bemx: ', what, ' failed with status ', code
    if (c_associated(bemx_ctx)) then
      p = bemx_last_error(bemx_ctx)
      call c_f_pointer(p, msg, [512])
      do k = 1, 512
        if (msg(k) == c_null_char) exit
        write(0, '(A)', advance='no') msg(k)
      end do
      write(0, *)
    end if
    stop 1
  end subroutine

  subroutine bemx_ensure_context()
    integer(c_int) :: code
    character(len=32) :: env
    integer :: dev, ios, elen
    if (c_associated(bemx_ctx)) return
    dev = 0
    call get_environment_variable('BEMX_DEVICE', env, elen)
    if (elen > 0) then
      read(env, *, iostat=ios) dev
      if (ios /= 0) dev = 0
    end if
    call get_environment_variable('BEMX_SOLVER', env, elen)
    if (elen > 0 .and. env(1:5) == 'mixed') bemx_solver_mode = BEMX_MODE_MIXED
    code = bemx_create(int(dev, c_int), bemx_ctx)
    call bemx_check(code, 'bemx_create')
  end subroutine

end module bemx_iface
