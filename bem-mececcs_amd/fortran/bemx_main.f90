! bemx_main.f90 -- driver with the reference's command line
!   bemx_main <E-deck> <D-deck> <E-out> <D-out>       (Main.for:85-101)
! It reads the decks (formats of Inputece.for:34-105 and Inputecd.for:26-134),
! runs the static pass once and, per frequency, GHMATECD -> LINSOLVE ->
! INTEREC through libbemx.so (the drop-in subroutines of bemx_surface.f90),
! and writes the reference's output files: the E output is INPUTECE's echo of
! the E deck (Inputece.for:29-104), the D output INPUTECD's echo of the D deck
! (Inputecd.for:23-136) followed by one OUTPUTEC block per frequency
! (Outputec.for:17-65), with the reference's FORMATs (ISO-8859-1 labels), so
! the files compare with the reference's own (src/tests.sh:42-55: cmp for the
! E output, dtester for the D output).
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
  ! ISO-8859-1 letters of the reference's labels
  character(len=*), parameter :: NUM = 'N' // achar(218) // 'MERO', MOD = 'M' // achar(211) // &
    'DULO', FREQ = 'FREQ' // achar(220) // achar(202) // 'NCIA', NOS = 'N' // achar(211) // 'S', &
    COND = 'CONDI' // achar(199) // achar(213) // 'ES', DIR = 'DIRE' // achar(199) // &
    achar(195) // 'O', COD = 'C' // achar(211) // 'DIGO'
  character(len=*), parameter :: STARS = "(/' ',79('*')/)"
  character(len=*), parameter :: F_E200 = "(//'DADOS'//2X,'" // NUM // &
    " DE ELEMENTOS DA MALHA= ',I4/2X,'" // NUM // " DE ELEMENTOS DE CONTORNO= ',I4/2X,'" // &
    NUM // " DE PONTOS EXTREMOS DOS ELEMENTOS= ',I4/2X,'" // NUM // &
    " DE PONTOS DE GAUSS= ',I3/2X,'" // MOD // " DE CISALHAMENTO= ',D14.7/2X," // &
    "'COEFICIENTE DE POISSON= ',D14.7)"
  character(len=*), parameter :: F_D200 = "(//'DADOS'//2X,'" // NUM // &
    " DE ELEMENTOS DA MALHA= ',I4/2X,'" // NUM // " DE ELEMENTOS DE CONTORNO= ',I4/2X,'" // &
    NUM // " DE PONTOS EXTREMOS DOS ELEMENTOS= ',I4/2X,'" // NUM // &
    " DE PONTOS INTERNOS= ',I3/2X,'" // NUM // " DE PONTOS DE GAUSS= ',I3/2X,'" // NUM // &
    " DE " // FREQ // "S= ',I3/2X,'" // MOD // " DE CISALHAMENTO (REAL)= ',D14.7/2X," // &
    "'COEFICIENTE DE POISSON= ',D14.7/2X,'COEFICIENTE DE AMORTECIMENTO= ',D14.7/2X," // &
    "'DENSIDADE DE MASSA= ',D14.7)"
  character(len=*), parameter :: F_D230 = "(//'DADOS DERIVADOS'//2X,'" // MOD // &
    " DE CISALHAMENTO (COMPLEXO)= ',2(D14.7)/2X,'VELOCIDADE DA ONDA-S (COMPLEXO)= '," // &
    "2(D14.7)/2X,'VELOCIDADE DA ONDA-P (COMPLEXO)= ',2(D14.7))"
  character(len=*), parameter :: F_XYZ = "(//2X,'COORDENADAS DOS PONTOS EXTREMOS DOS " // &
    "ELEMENTOS ',A//4X,'PONTO',10X,'X',18X,'Y',18X,'Z')"
  character(len=*), parameter :: F_CONE = "(//2X,'CONECTIVIDADE DOS ELEMENTOS ',A//1X," // &
    "'ELEMENTO',3X,'EXTR I',3X,'EXTR J',3X,'EXTR K',3X,'EXTR L',5X,'XMED',8X,'YMED',8X,'ZMED')"
  character(len=*), parameter :: F_D600 = "(//2X,'" // COND // " DE CONTORNO'//8X," // &
    "3('VALOR PRESCRITO',10X)/1X,'ELEM',6X,'" // DIR // " X',4X,'" // COD // "',6X,'" // &
    DIR // " Y',4X,'" // COD // "',6X,'" // DIR // " Z',4X,'" // COD // "')"
  character(len=*), parameter :: F_D700 = "(//2X,'COORDENADAS DOS PONTOS INTERNOS '," // &
    "//4X,'PONTO',10X,'X',18X,'Y',17X,' Z')"
  character(len=*), parameter :: F_O700 = "(30X,'" // FREQ // " = ',D14.7/)"
  character(len=*), parameter :: F_O100 = "(' ',79('*')//1X,'RESULTADOS'//2X,'" // NOS // &
    " DO CONTORNO'//' ELEM',6X,'DESLOCAMENTO X',11X,'DESLOCAMENTOT Y',11X," // &
    "'DESLOCAMENTO Z'/)"
  character(len=*), parameter :: F_O150 = "(/' ELEM',8X,'TRACTION X',15X,'TRACTION Y',15X," // &
    "'TRACTION Z'/)"
  character(len=*), parameter :: F_O300 = "(//2X,'PONTOS INTERNOS'//' PONTO',5X," // &
    "'DESLOCAMENTO X',11X,'DESLOCAMENTO Y',11X,'DESLOCAMENTO Z')"
  character(len=*), parameter :: F_O400 = "(//2X,'PONTOS INTERNOS'//' PONTO',8X,'SIGMAXX'," // &
    "17X,'SIGMAXY',18X,'SIGMAXZ',18X,'SIGMAYX',18X,'SIGMAYY',18X,'SIGMAYZ',18X,'SIGMAZX'," // &
    "18X,'SIGMAZY',18X,'SIGMAZZ')"
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
  ! INPUTECE's echo (Inputece.for:29-104)
  write(IPR, STARS)
  write(IPR, '(A)') title
  write(IPR, F_E200) n, nbe, np, npg, ge, rnu
  write(IPR, F_XYZ) 'DE CONTORNO'
  do i = 1, np
    write(IPR, 400) i, cx(i), cy(i), cz(i)
  end do
  write(IPR, F_CONE) 'DE CONTORNO'
  do i = 1, n
    write(IPR, 550) i, (cone(i,j), j = 1, 4), cxm(i), cym(i), czm(i)
  end do

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
  ! INPUTECD's echo (Inputecd.for:23-136)
  write(IPS, STARS)
  write(IPS, '(A)') title
  write(IPS, F_D200) n, nbe, np, l, npg, nfr, ge, rnu, dam, rho
  write(IPS, F_D230) zge, zcs, zcp
  write(IPS, "(/2X,'FREQUENCIAS'//5(2X,D14.7))") (afr(i), i = 1, nfr)
  write(IPS, F_XYZ) 'DA MALHA'
  do i = 1, np
    write(IPS, 400) i, cx(i), cy(i), cz(i)
  end do
  write(IPS, F_CONE) 'DA MALHA'
  do i = 1, n
    write(IPS, 550) i, (cone(i,j), j = 1, 4), cxm(i), cym(i), czm(i)
  end do
  write(IPS, F_D600)
  do i = 1, nbe
    write(IPS, 650) i, (bc(3*(i-1)+j), kode(3*(i-1)+j), j = 1, 3)
  end do
  if (l > 0) then
    write(IPS, F_D700)
    write(IPS, 750) (i, cxi(i), cyi(i), czi(i), i = 1, l)
  end if

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
    ! ---- OUTPUTEC (Outputec.for:17-65)
    write(IPS, STARS)
    write(IPS, F_O700) fr
    write(IPS, F_O100)
    do i = 1, nbe
      write(IPS, 200) i, (zfi(3*(i-1)+j), j = 1, 3)
    end do
    write(IPS, F_O150)
    do i = 1, nbe
      write(IPS, 200) i, (zdfi(3*(i-1)+j), j = 1, 3)
    end do
    if (l > 0) then
      write(IPS, F_O300)
      do i = 1, l
        write(IPS, 200) i, (zdsol(3*(i-1)+j), j = 1, 3)
      end do
      write(IPS, F_O400)
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
400 format(4X, I4, 3(5X, D14.7))
550 format(I6, 4I9, 2X, 3D12.4)
650 format(I4, 3(4X, D10.3, 6X, I3, 2X))
750 format(4X, I3, 3D19.7)

  call DEALLOCATE_SHARED_GPU_DATA()
  code = bemx_destroy(bemx_ctx)
  close(INP); close(INQ); close(IPR); close(IPS)
end program bemx_main
