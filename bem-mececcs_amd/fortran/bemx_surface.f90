! bemx_surface.f90 -- the reference's Fortran subroutine surface, served by
! libbemx.so.  Same names, argument lists, types (REAL = real(8), the
! FPREC=double build) and allocation conventions as the reference, so its
! driver links against these in place of its own routines:
!   GHMATECE    Ghmatece.for:9-11   (interface Ghmatece.fd:2-18)
!   GHMATECD    Ghmatecd.for:11-13  (interface Ghmatecd.fd:2-23)
!   LINSOLVE    Linsolve.for:1      (interface Linsolve.fd:3-8)
!   INTEREC     Interec.for:9-11    (interface Interec.fd:3-19)
! and the device-data entry points of the reference's GPU build:
!   SEND_SHARED_DATA_TO_GPU   kernels/shared.fd:3-14
!   DEALLOCATE_SHARED_GPU_DATA kernels/shared.fd:16-18
!   SEND_DYN_DATA_GPU          (implicit external, Main.for:162)
! Environment: BEMX_DEVICE selects the GPU, BEMX_SOLVER=mixed selects the
! mixed-precision LINSOLVE (FP32 BF16x3 LU + FP64 GMRES-IR), default FP64.

subroutine SEND_SHARED_DATA_TO_GPU(CX, CY, CZ, CXM, CYM, CZM, ETAS, GI, OME, CONE, NP, NPG, &
                                   N, NBE)
  use bemx_iface
  implicit none
  integer, intent(in) :: NP, NPG, N, NBE
  real(8), intent(in) :: CX(NP), CY(NP), CZ(NP), CXM(N), CYM(N), CZM(N), ETAS(3,N)
  real(8), intent(in) :: GI(NPG), OME(NPG)
  integer, intent(in) :: CONE(N,4)
  integer(c_int) :: code
  real(8) :: cxv(NP), cyv(NP), czv(NP), cxmv(N), cymv(N), czmv(N), etv(3,N), giv(NPG), omv(NPG)
  integer(c_int) :: conev(N,4)
  call bemx_ensure_context()
  cxv = CX; cyv = CY; czv = CZ; cxmv = CXM; cymv = CYM; czmv = CZM; etv = ETAS
  giv = GI; omv = OME; conev = CONE
  code = bemx_send_shared_data(bemx_ctx, cxv, cyv, czv, cxmv, cymv, czmv, etv, giv, omv, &
                               conev, int(NP, c_int), int(NPG, c_int), int(N, c_int), &
                               int(NBE, c_int))
  call bemx_check(code, 'SEND_SHARED_DATA_TO_GPU')
  bemx_mesh_sent = .true.
end subroutine

subroutine DEALLOCATE_SHARED_GPU_DATA()
  use bemx_iface
  implicit none
  integer(c_int) :: code
  if (.not. c_associated(bemx_ctx)) return
  code = bemx_deallocate_shared_data(bemx_ctx)
  call bemx_check(code, 'DEALLOCATE_SHARED_GPU_DATA')
  bemx_mesh_sent = .false.
end subroutine

subroutine SEND_DYN_DATA_GPU(NBE, ZDFI, KODE)
  use bemx_iface
  implicit none
  integer, intent(in) :: NBE
  complex(8), intent(in) :: ZDFI(3*NBE)
  integer, intent(in) :: KODE(3*NBE)
  complex(c_double_complex) :: z(3*NBE)
  integer(c_int) :: k(3*NBE), code
  call bemx_ensure_context()
  z = ZDFI
  k = KODE
  code = bemx_send_dyn_data(bemx_ctx, z, k)
  call bemx_check(code, 'SEND_DYN_DATA_GPU')
end subroutine

! Static pass: HESTdiag/GESTdiag (allocated here, as the reference does) and
! the constants PI, C1..C4, DELTA returned to the caller (Ghmatece.for:55-72).
subroutine GHMATECE(CX, CY, CZ, CXM, CYM, CZM, HESTdiag, GESTdiag, NE, NX, NCOX, CONE, N, &
                    NBE, NP, NPG, GE, RNU, RMU, DELTA, PI, C1, C2, C3, C4, ETAS, GI, OME)
  use bemx_iface
  implicit none
  integer, intent(in) :: NE, NX, NCOX, N, NBE, NP, NPG
  real(8), intent(in) :: CX(NP), CY(NP), CZ(NP), CXM(N), CYM(N), CZM(N)
  real(8), dimension(:,:,:), allocatable, intent(out) :: HESTdiag, GESTdiag
  integer, intent(in) :: CONE(N,4)
  real(8), intent(in) :: GE, RNU, RMU
  real(8), intent(out) :: DELTA(3,3), PI, C1, C2, C3, C4
  real(8), intent(in) :: ETAS(3,N), GI(NPG), OME(NPG)
  integer(c_int) :: code
  real(8) :: hd(3,3,NBE), gd(3,3,NBE), dl(3,3), p, a1, a2, a3, a4
  if (.not. bemx_mesh_sent) &
    call SEND_SHARED_DATA_TO_GPU(CX, CY, CZ, CXM, CYM, CZM, ETAS, GI, OME, CONE, NP, NPG, N, NBE)
  code = bemx_ghmatece(bemx_ctx, RMU, RNU, hd, gd, dl, p, a1, a2, a3, a4)
  call bemx_check(code, 'GHMATECE')
  allocate(HESTdiag(3,3,NBE), GESTdiag(3,3,NBE))
  HESTdiag = hd
  GESTdiag = gd
  DELTA = dl
  PI = p; C1 = a1; C2 = a2; C3 = a3; C4 = a4
end subroutine

! Dynamic assembly at frequency FR: ZH (3NBE x 3NBE, after the KODE column
! swap) and ZFI = ZG ZDFI, both allocated here (Ghmatecd.for:82-92).  The
! matrix also stays on the device for the LINSOLVE that follows.
subroutine GHMATECD(CX, CY, CZ, CXM, CYM, CZM, HESTdiag, GESTdiag, ZH, KODE, NE, NX, NCOX, &
                    CONE, DELTA, PI, N, NBE, NP, NPG, GE, RNU, RMU, ZDFI, ZFI, L, FR, DAM, &
                    RHO, ZGE, ZCS, ZCP, C1, C2, C3, C4, ETAS, GI, OME)
  use bemx_iface
  implicit none
  integer, intent(in) :: NE, NX, NCOX, N, NBE, NP, NPG, L
  real(8), intent(in) :: CX(NP), CY(NP), CZ(NP), CXM(N), CYM(N), CZM(N)
  real(8), intent(in) :: HESTdiag(3,3,NBE), GESTdiag(3,3,NBE)
  complex(8), dimension(:,:), allocatable, intent(out) :: ZH
  integer, intent(in) :: KODE(3*NBE), CONE(N,4)
  real(8), intent(in) :: DELTA(3,3), PI, GE, RNU, RMU, FR, DAM, RHO, C1, C2, C3, C4
  complex(8), intent(in) :: ZGE, ZCS, ZCP, ZDFI(3*NBE)
  complex(8), allocatable, intent(out) :: ZFI(:)
  real(8), intent(in) :: ETAS(3,N), GI(NPG), OME(NPG)
  integer(c_int) :: code
  real(8) :: hd(3,3,NBE), gd(3,3,NBE)
  complex(c_double_complex) :: zg, zs, zp
  if (.not. bemx_mesh_sent) &
    call SEND_SHARED_DATA_TO_GPU(CX, CY, CZ, CXM, CYM, CZM, ETAS, GI, OME, CONE, NP, NPG, N, NBE)
  call SEND_DYN_DATA_GPU(NBE, ZDFI, KODE)
  allocate(ZFI(3*NBE), ZH(3*NBE, 3*NBE))
  hd = HESTdiag
  gd = GESTdiag
  zg = ZGE; zs = ZCS; zp = ZCP
  code = bemx_ghmatecd(bemx_ctx, zg, zs, zp, C1, C2, C3, C4, FR, hd, gd, ZFI, ZH)
  call bemx_check(code, 'GHMATECD')
end subroutine

! ZH x = ZFI; x returned in ZFI, ZH destroyed (Linsolve.for:1-66).  A zero
! pivot is reported and the run continues, as LINSOLVE_CPU does
! (Linsolve.for:102-106).
subroutine LINSOLVE(NN, N, ZH, ZFI)
  use bemx_iface
  implicit none
  integer, intent(in) :: NN, N
  complex(8), intent(inout) :: ZH(NN, NN)
  complex(8), intent(inout) :: ZFI(NN)
  integer(c_int) :: code
  integer(c_long_long) :: info
  call bemx_ensure_context()
  code = bemx_linsolve(bemx_ctx, int(NN, c_int), ZH, ZFI, bemx_solver_mode, info)
  if (code == 2) then
    print *, "Matriz Singular :-|"
    return
  end if
  call bemx_check(code, 'LINSOLVE')
end subroutine

! Interior points: un-swap ZFI/ZDFI in place, then ZDSOL(3L) and ZSSOL(9L)
! (allocated here; nothing is allocated when L = 0, as Interec.for:71).
subroutine INTEREC(ZFI, ZDFI, KODE, CX, CY, CZ, CXI, CYI, CZI, ZDSOL, ZSSOL, NE, NX, NCOX, &
                   NPIX, CONE, ZGE, ZCS, ZCP, DELTA, PI, FR, NPG, L, N, NBE, RHO, ETAS, GI, &
                   OME, NP)
  use bemx_iface
  implicit none
  integer, intent(in) :: NE, NX, NCOX, NPIX, NPG, L, N, NBE, NP
  complex(8), dimension(3*NBE), intent(inout) :: ZFI, ZDFI
  integer, intent(in) :: KODE(3*NBE), CONE(N,4)
  real(8), dimension(:), intent(in) :: CX, CY, CZ
  real(8), dimension(L), intent(in) :: CXI, CYI, CZI
  complex(8), dimension(:), allocatable, intent(out) :: ZDSOL, ZSSOL
  complex(8), intent(in) :: ZGE, ZCS, ZCP
  real(8), intent(in) :: DELTA(3,3), PI, FR, RHO
  real(8), intent(in) :: ETAS(3,NX), GI(NPG), OME(NPG)
  integer(c_int) :: code, k(3*NBE)
  real(8) :: xi(max(L,1)), yi(max(L,1)), zi(max(L,1))
  complex(c_double_complex) :: zg, zs, zp
  complex(c_double_complex), allocatable :: ds(:), ss(:)
  call bemx_ensure_context()
  k = KODE
  zg = ZGE; zs = ZCS; zp = ZCP
  if (L > 0) then
    xi(1:L) = CXI; yi(1:L) = CYI; zi(1:L) = CZI
  end if
  allocate(ds(max(3*L,1)), ss(max(9*L,1)))
  code = bemx_interec(bemx_ctx, ZFI, ZDFI, k, int(L, c_int), xi, yi, zi, zg, zs, zp, FR, RHO, &
                      ds, ss)
  call bemx_check(code, 'INTEREC')
  if (L <= 0) return
  allocate(ZDSOL(3*L), ZSSOL(9*L))
  ZDSOL = ds(1:3*L)
  ZSSOL = ss(1:9*L)
end subroutine
