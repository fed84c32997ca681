! bemx_main.f90 -- driver with the reference's command line
!   bemx_main <E-deck> <D-deck> <E-out> <D-out>       (Main.for:85-101)
! It reads the decks (formats of Inputece.for:34-105 and Inputecd.for:26-134),
! runs the static pass once and, per frequency, GHMATECD -> LINSOLVE ->
! INTEREC through libbemx.so (the drop-in subroutines of bemx_surface.f90),
! and writes the results in the reference's output layout (Outputec.for:9-68:
! FORMAT I5,3(2X,D11.4,',',D11.4) for boundary and interior displacements,
! I4,9(...) for the interior stresses).
! Deviation from the reference: ZDFI is reset to the prescribed values at
! every frequency (the reference sets it once, Main.for:158, and INTEREC's
! in-place un-swap then corrupts later frequencies when some KODE = 0).
program bemx_main
  use bemx_iface
  implicit none
  interface
    subroutine GHMATECE(CX, CY, CZ, CXM, CYM, CZM, HESTdiag, GESTdiag, NE, NX, NCOX, CONE, N, &
                        NBE, NP, NPG, GE, RNU, RMU, DELTA, PI, C1, C2, C3, C4, ETAS, GI, OME)
      integer, intent(in) :: NE, NX, NCOX, N, NBE, NP, NPG
      real(8), intent(in) :: CX(NP), CY(NP), CZ(NP), CXM(N), CYM(N), CZM(N)
      real(8), dimension(:,:,:), allocatable, intent(out) :: HESTdiag, GESTdiag
      integer, intent(in) :: CONE(N,4)
      real(8), intent(in) :: GE, RNU, RMU
      real(8), intent(out) :: DELTA(3,3), PI, C1, C2, C3, C4
      real(8), intent(in) :: ETAS(3,N), GI(NPG), OME(NPG)
    end subroutine
    subroutine GHMATECD(CX, CY, CZ, CXM, CYM, CZM, HESTdiag, GESTdiag, ZH, KODE, NE, NX, NCOX, &
                        CONE, DELTA, PI, N, NBE, NP, NPG, GE, RNU, RMU, ZDFI, ZFI, L, FR, DAM, &
                        RHO, ZGE, ZCS, ZCP, C1, C2, C3, C4, ETAS, GI, OME)
      integer, intent(in) :: NE, NX, NCOX, N, NBE, NP, NPG, L
      real(8), intent(in) :: CX(NP), CY(NP), CZ(NP), CXM(N), CYM(N), CZM(N)
      real(8), intent(in) :: HESTdiag(3,3,NBE), GESTdiag(3,3,NBE)
      complex(8), dimension(:,:), allocatable, intent(out) :: ZH
      integer, intent(in) :: KODE(3*NBE), CONE(N,4)
      real(8), intent(in) :: DELTA(3,3), PI, GE, RNU, RMU, FR, DAM, RHO, C1, C2, C3, C4
      complex(8), intent(in) :: ZGE, ZCS, ZCP, ZDFI(3*NBE)
      complex(8), allocatable, intent(out) :: ZFI(:)
      real(8), intent(in) :: ETAS(3,N), GI(NPG), OME(NPG)
    end subroutine
    subroutine LINSOLVE(NN, N, ZH, ZFI)
      integer, intent(in) :: NN, N
      complex(8), intent(inout) :: ZH(NN, NN), ZFI(NN)
    end subroutine
    subroutine INTEREC(ZFI, ZDFI, KODE, CX, CY, CZ, CXI, CYI, CZI, ZDSOL, ZSSOL, NE, NX, NCOX, &
                       NPIX, CONE, ZGE, ZCS, ZCP, DELTA, PI, FR, NPG, L, N, NBE, RHO, ETAS, GI, &
                       OME, NP)
      integer, intent(in) :: NE, NX, NCOX, NPIX, NPG, L, N, NBE, NP
      complex(8), dimension(3*NBE), intent(inout) :: ZFI, ZDFI
      integer, intent(in) :: KODE(3*NBE), CONE(N,4)
      real(8), dimension(:), intent(in) :: CX, CY, CZ
      real(8), dimension(L), intent(in) :: CXI, CYI, CZI
      complex(8), dimension(:), allocatable, intent(out) :: ZDSOL, ZSSOL
      complex(8), intent(in) :: ZGE, ZCS, ZCP
      real(8), intent(in) :: DELTA(3,3), PI, FR, RHO
      real(8), intent(in) :: ETAS(3,NX), GI(NPG), OME(NPG)
    end subroutine
  end interface

  integer, parameter :: INP = 100, INQ = 101, IPR = 102, IPS = 103, IPT = 104
  character(len=256) :: fe, fd, foe, fod
  character(len=75) :: title
  integer :: n, nbe, np, npg, l, nfr, i, j, k, ifr, nn
  real(8) :: ge, rnu, rmu, dam, rho, fr, pi, c1, c2, c3, c4, delta(3,3), a, b, c, r, conein(4)
  real(8), allocatable :: cx(:), cy(:), cz(:), cxm(:), cym(:), czm(:), etas(:,:), gi(:), ome(:)
  real(8), allocatable :: cxi(:), cyi(:), czi(:), bc(:), afr(:), hd(:,:,:), gd(:,:,:)
  integer, allocatable :: cone(:,:), kode(:)
  complex(8), allocatable :: zh(:,:), zfi(:), zdfi(:), zdsol(:), zssol(:)
  complex(8) :: zge, zcs, zcp
  integer(c_int) :: code

  if (command_argument_count() < 4) then
    write(0, '(A)') 'usage: bemx_main <E-deck> <D-deck> <E-out> <D-out>'
    stop 2
  end if
  call get_command_argument(1, fe)
  call get_command_argument(2, fd)
  call get_command_argument(3, foe)
  call get_command_argument(4, fod)
  open(INP, file=fe, status='old')
  open(INQ, file=fd, status='old')
  open(IPR, file=foe, status='unknown')
  open(IPS, file=fod, status='unknown')

  ! ---- E deck
  read(INP, '(A)') title
  read(INP, *) n, nbe, np, npg, ge, rnu
  rmu = ge
  allocate(cx(np), cy(np), cz(np), cxm(n), cym(n), czm(n), cone(n,4), etas(3,n))
  read(INP, *) (cx(i), cy(i), cz(i), i = 1, np)
  do i = 1, n
    read(INP, *) (conein(j), j = 1, 4)
    cone(i, 1:4) = int(conein(1:4))
    cxm(i) = (cx(cone(i,1)) + cx(cone(i,2)) + cx(cone(i,3)) + cx(cone(i,4))) * 0.25d0
    cym(i) = (cy(cone(i,1)) + cy(cone(i,2)) + cy(cone(i,3)) + cy(cone(i,4))) * 0.25d0
    czm(i) = (cz(cone(i,1)) + cz(cone(i,2)) + cz(cone(i,3)) + cz(cone(i,4))) * 0.25d0
  end do
  write(IPR, '(A)') title
  write(IPR, '(A,I6,A,I6,A,I6,A,I3,A,D14.7,A,D14.7)') ' N=', n, ' NBE=', nbe, ' NP=', np, &
    ' NPG=', npg, ' G=', ge, ' NU=', rnu

  ! Gauss rule and unit normals (P2-P1) x (P3-P1)
  allocate(gi(npg), ome(npg))
  code = bemx_gauleg(int(npg, c_int), gi, ome)
  do j = 1, n
    a = (cy(cone(j,2)) - cy(cone(j,1))) * (cz(cone(j,3)) - cz(cone(j,1))) - &
        (cz(cone(j,2)) - cz(cone(j,1))) * (cy(cone(j,3)) - cy(cone(j,1)))
    b = (cz(cone(j,2)) - cz(cone(j,1))) * (cx(cone(j,3)) - cx(cone(j,1))) - &
        (cx(cone(j,2)) - cx(cone(j,1))) * (cz(cone(j,3)) - cz(cone(j,1)))
    c = (cx(cone(j,2)) - cx(cone(j,1))) * (cy(cone(j,3)) - cy(cone(j,1))) - &
        (cy(cone(j,2)) - cy(cone(j,1))) * (cx(cone(j,3)) - cx(cone(j,1)))
    r = sqrt(a*a + b*b + c*c)
    etas(1,j) = a / r; etas(2,j) = b / r; etas(3,j) = c / r
  end do

  call GHMATECE(cx, cy, cz, cxm, cym, czm, hd, gd, n, 3*n, np, cone, n, nbe, np, npg, ge, rnu, &
                rmu, delta, pi, c1, c2, c3, c4, etas, gi, ome)

  ! ---- D deck
  read(INQ, '(A)') title
  read(INQ, *) l, nfr, dam, rho
  allocate(afr(nfr), kode(3*nbe), bc(3*nbe), cxi(max(l,1)), cyi(max(l,1)), czi(max(l,1)))
  zge = cmplx(ge, ge*2.0d0*dam, kind=8)
  zcs = sqrt(zge / rho)
  zcp = zcs * sqrt((2.0d0 - 2.0d0*rnu) / (1.0d0 - 2.0d0*rnu))
  read(INQ, *) (afr(i), i = 1, nfr)
  do i = 1, nbe
    read(INQ, *) (kode(3*(i-1)+j), bc(3*(i-1)+j), j = 1, 3)
  end do
  if (l > 0) read(INQ, *) (cxi(i), cyi(i), czi(i), i = 1, l)
  write(IPS, '(A)') title
  write(IPS, '(A,I6,A,I6,A,I6,A,I4,A,I3,A,I4)') ' N=', n, ' NBE=', nbe, ' NP=', np, ' L=', l, &
    ' NPG=', npg, ' NFR=', nfr

  nn = 3*nbe
  allocate(zdfi(nn))
  do ifr = 1, nfr
    fr = afr(ifr)
    zdfi = bc
    call GHMATECD(cx, cy, cz, cxm, cym, czm, hd, gd, zh, kode, n, 3*n, np, cone, delta, pi, n, &
                  nbe, np, npg, ge, rnu, rmu, zdfi, zfi, l, fr, dam, rho, zge, zcs, zcp, c1, &
                  c2, c3, c4, etas, gi, ome)
    call LINSOLVE(nn, n, zh, zfi)
    deallocate(zh)
    call INTEREC(zfi, zdfi, kode, cx, cy, cz, cxi(1:l), cyi(1:l), czi(1:l), zdsol, zssol, n, n, &
                 np, l, cone, zge, zcs, zcp, delta, pi, fr, npg, l, n, nbe, rho, etas, gi, ome, np)
    ! ---- results block
    write(IPS, '(/1X,79(''*'')/)')
    write(IPS, '(30X,A,D14.7/)') 'FREQUENCIA = ', fr
    write(IPS, '(1X,79(''*'')//1X,A//2X,A//1X,A/)') 'RESULTADOS', 'NOS DO CONTORNO', &
      'ELEM      DESLOCAMENTO X           DESLOCAMENTO Y           DESLOCAMENTO Z'
    do i = 1, nbe
      write(IPS, 200) i, (zfi(3*(i-1)+j), j = 1, 3)
    end do
    write(IPS, '(/1X,A/)') 'ELEM        TRACTION X               TRACTION Y               TRACTION Z'
    do i = 1, nbe
      write(IPS, 200) i, (zdfi(3*(i-1)+j), j = 1, 3)
    end do
    if (l > 0) then
      write(IPS, '(//2X,A//1X,A)') 'PONTOS INTERNOS', &
        'PONTO     DESLOCAMENTO X           DESLOCAMENTO Y           DESLOCAMENTO Z'
      do i = 1, l
        write(IPS, 200) i, (zdsol(3*(i-1)+j), j = 1, 3)
      end do
      write(IPS, '(//2X,A//1X,A)') 'PONTOS INTERNOS', &
        'PONTO        SIGMAXX                 SIGMAXY                  SIGMAXZ'
      do i = 1, l
        write(IPS, 350) i, (zssol(9*(i-1)+k), k = 1, 9)
      end do
      if (nn >= 166) then
        open(IPT, file='fort.104', status='unknown', position='append')
        write(IPT, *) fr, abs(zfi(157)), abs(zfi(166))
        close(IPT)
      end if
      deallocate(zdsol, zssol)
    end if
    write(IPS, '(1X,230(''*''))')
    deallocate(zfi)
  end do
200 format(I5, 3(2X, D11.4, ',', D11.4))
350 format(I4, 9(2X, D11.4, ',', D11.4))

  call DEALLOCATE_SHARED_GPU_DATA()
  code = bemx_destroy(bemx_ctx)
  close(INP); close(INQ); close(IPR); close(IPS)
end program bemx_main
