! ref_harness.f90 -- C-callable harness around the reference's own Fortran
! routines, compiled from the reference sources by oracle/Makefile into
! oracle/_ref/libbemref.so.
!
! TEST INFRASTRUCTURE ONLY: used to pin the C restatement (bem_oracle.c)
! against the reference itself.  The reference's GHMATECD and LINSOLVE
! subroutines call BLAS/LAPACK (ZGEMV, ZGESV) and are not compiled here; the
! dynamic assembly loop of GHMATECD (Ghmatecd.for:115-165) is restated below
! around the reference's own SING_DE and NONSINGD, and the solve is pinned
! separately against LAPACK through numpy.
!
! Everything here is compiled with the reference's FPREC=double flags
! (-fdefault-real-8 -fdefault-double-8), so REAL == real(8) for the
! reference routines called through implicit interfaces.
module bemref_harness
  use iso_c_binding
  implicit none

  interface
    subroutine GHMATECE(CX,CY,CZ,CXM,CYM,CZM,HESTdiag,GESTdiag,NE,NX,NCOX, &
                        CONE,N,NBE,NP,NPG,GE,RNU,RMU,DELTA,PI,C1,C2,C3,C4,ETAS,GI,OME)
      integer :: NE, NX, NCOX, N, NBE, NP, NPG
      real(8), dimension(NP) :: CX, CY, CZ
      real(8), dimension(N) :: CXM, CYM, CZM
      real(8), dimension(:,:,:), allocatable, intent(out) :: HESTdiag, GESTdiag
      integer, dimension(N,4) :: CONE
      real(8) :: GE, RNU, RMU, DELTA(3,3), PI, C1, C2, C3, C4, ETAS(3,N)
      real(8), dimension(NPG) :: GI, OME
    end subroutine
    subroutine INTEREC(ZFI,ZDFI,KODE,CX,CY,CZ,CXI,CYI,CZI,ZDSOL,ZSSOL,NE,NX,NCOX, &
                       NPIX,CONE,ZGE,ZCS,ZCP,DELTA,PI,FR,NPG,L,N,NBE,RHO,ETAS,GI,OME,NP)
      integer :: NE, NX, NCOX, NPIX, NPG, L, N, NBE, NP
      real(8), dimension(:) :: CX, CY, CZ
      real(8), dimension(L) :: CXI, CYI, CZI
      real(8) :: DELTA(3,3), PI, FR, RHO
      complex(8), dimension(3*NBE) :: ZDFI, ZFI
      complex(8), allocatable, intent(out) :: ZDSOL(:), ZSSOL(:)
      integer :: CONE(N,4), KODE(3*NBE)
      complex(8) :: ZGE, ZCS, ZCP
      real(8) :: ETAS(3,NX)
      real(8), dimension(NPG) :: GI, OME
    end subroutine
    function NORMVEC(cone, cx, cy, cz, n, np) result(ETAS)
      integer :: n, np
      integer, dimension(n,4) :: cone
      real(8), dimension(np) :: cx, cy, cz
      real(8), dimension(:,:), allocatable :: ETAS
    end function
  end interface

  integer :: INP, INQ, IPR, IPS, IPT
  common INP, INQ, IPR, IPS, IPT

contains

  subroutine set_units()
    INP = 100; INQ = 101; IPR = 6; IPS = 6; IPT = 6
  end subroutine

  subroutine identity3(d)
    real(8), intent(out) :: d(3,3)
    integer :: i, j
    do j = 1, 3
      do i = 1, 3
        d(i,j) = merge(1.0d0, 0.0d0, i == j)
      end do
    end do
  end subroutine

  subroutine ref_gauleg(npg, gi, ome) bind(C, name='ref_gauleg')
    integer(c_int), value :: npg
    real(c_double) :: gi(npg), ome(npg)
    call GAULEG(-1.0d0, 1.0d0, gi, ome, npg)
  end subroutine

  subroutine ref_normvec(n, np, cone, cx, cy, cz, etas) bind(C, name='ref_normvec')
    integer(c_int), value :: n, np
    integer(c_int) :: cone(n,4)
    real(c_double) :: cx(np), cy(np), cz(np), etas(3,n)
    etas = NORMVEC(cone, cx, cy, cz, n, np)
  end subroutine

  ! Material constants exactly as INPUTECD computes them (Inputecd.for:58-60).
  subroutine ref_material(ge, rnu, dam, rho, z) bind(C, name='ref_material')
    real(c_double), value :: ge, rnu, dam, rho
    complex(c_double_complex) :: z(3)
    z(1) = cmplx(ge, ge*2.0d0*dam, kind=8)
    z(2) = sqrt(z(1)/rho)
    z(3) = z(2)*sqrt((2.0d0-2.0d0*rnu)/(1.0d0-2.0d0*rnu))
  end subroutine

  subroutine ref_solfund(xp, xg, rn, zge, zcs, zcp, pi, fr, zu, zt) bind(C, name='ref_solfund')
    real(c_double) :: xp(3), xg(3), rn(3)
    complex(c_double_complex), value :: zge, zcs, zcp
    real(c_double), value :: pi, fr
    complex(c_double_complex) :: zu(3,3), zt(3,3)
    real(8) :: delta(3,3)
    call identity3(delta)
    call SOLFUND(zu, zt, xp(1), xp(2), xp(3), xg(1), xg(2), xg(3), rn, zge, zcs, zcp, &
                 delta, pi, fr)
  end subroutine

  subroutine ref_solfundif(xp, xg, rn, zm, c1, c2, c3, c4, pi, fr, zu, zt) &
      bind(C, name='ref_solfundif')
    real(c_double) :: xp(3), xg(3), rn(3)
    complex(c_double_complex) :: zm(3)
    real(c_double), value :: c1, c2, c3, c4, pi, fr
    complex(c_double_complex) :: zu(3,3), zt(3,3)
    real(8) :: delta(3,3)
    call identity3(delta)
    call SOLFUNDIF(zu, zt, xp(1), xp(2), xp(3), xg(1), xg(2), xg(3), rn, zm(1), zm(2), zm(3), &
                   c1, c2, c3, c4, delta, pi, fr)
  end subroutine

  subroutine ref_solfune(xp, xg, eta, c3, c4, t) bind(C, name='ref_solfune')
    real(c_double) :: xp(3), xg(3), eta(3), t(3,3)
    real(c_double), value :: c3, c4
    real(8) :: delta(3,3)
    call identity3(delta)
    call SOLFUNE(t, xp(1), xp(2), xp(3), xg(1), xg(2), xg(3), eta, 0, 0, 0, 0.0d0, 0.0d0, &
                 c3, c4, delta)
  end subroutine

  subroutine ref_nonsinge(co, xp, eta, c3, c4, npg, gi, ome, h) bind(C, name='ref_nonsinge')
    real(c_double) :: co(4,3), xp(3), eta(3), h(3,3)
    real(c_double), value :: c3, c4
    integer(c_int), value :: npg
    real(c_double) :: gi(npg), ome(npg)
    real(8) :: delta(3,3)
    call set_units()
    call identity3(delta)
    call NONSINGE(h, co, xp(1), xp(2), xp(3), eta, 0, 0, npg, 0.0d0, 0.0d0, c3, c4, &
                  delta, gi, ome)
  end subroutine

  subroutine ref_singge(corners, xm, c1, c2, npg, gi, ome, g) bind(C, name='ref_singge')
    real(c_double) :: corners(4,3), xm(3), g(3,3)
    real(c_double), value :: c1, c2
    integer(c_int), value :: npg
    real(c_double) :: gi(npg), ome(npg)
    real(8) :: delta(3,3), cx(4), cy(4), cz(4), eta(3)
    call set_units()
    call identity3(delta)
    cx = corners(:,1); cy = corners(:,2); cz = corners(:,3)
    eta = 0.0d0
    call SINGGE(g, xm(1), xm(2), xm(3), eta, cx, cy, cz, 1, 2, 3, 4, 4, 1, 4, npg, &
                c1, c2, delta, gi, ome)
  end subroutine

  subroutine ref_nonsingd(co, xp, eta, zge, zcs, zcp, pi, fr, npg, gi, ome, zh, zg) &
      bind(C, name='ref_nonsingd')
    real(c_double) :: co(4,3), xp(3), eta(3)
    complex(c_double_complex), value :: zge, zcs, zcp
    real(c_double), value :: pi, fr
    integer(c_int), value :: npg
    real(c_double) :: gi(npg), ome(npg)
    complex(c_double_complex) :: zh(3,3), zg(3,3)
    real(8) :: delta(3,3)
    call set_units()
    call identity3(delta)
    call NONSINGD(zh, zg, co, xp(1), xp(2), xp(3), eta, zge, zcs, zcp, delta, pi, fr, &
                  gi, ome, npg)
  end subroutine

  ! The material (zge, zcs, zcp) comes by reference: with more than eight
  ! floating-point arguments a by-value complex would straddle registers and
  ! stack, which C and Fortran compilers need not lay out alike.
  subroutine ref_sigmaec(co, xp, rn, pi, fr, zm, rho, npg, gi, ome, zd, zs) &
      bind(C, name='ref_sigmaec')
    real(c_double) :: co(4,3), xp(3), rn(3)
    real(c_double), value :: pi, fr, rho
    complex(c_double_complex) :: zm(3)
    integer(c_int), value :: npg
    real(c_double) :: gi(npg), ome(npg)
    complex(c_double_complex) :: zd(3,3,3), zs(3,3,3)
    real(8) :: delta(3,3)
    call set_units()
    call identity3(delta)
    call SIGMAEC(co, xp(1), xp(2), xp(3), rn, delta, pi, fr, zm(1), rho, zm(2), zm(3), npg, &
                 zd, zs, gi, ome)
  end subroutine

  ! Geometry exactly as the reference's driver prepares it: Gauss rule
  ! (Main.for:112-114), normals (Main.for:120), centroids (Inputece.for:100-102).
  subroutine mesh_geometry(n, np, npg, cone, cx, cy, cz, cxm, cym, czm, etas, gi, ome)
    integer, intent(in) :: n, np, npg
    integer, intent(in) :: cone(n,4)
    real(8), intent(in) :: cx(np), cy(np), cz(np)
    real(8), intent(out) :: cxm(n), cym(n), czm(n), etas(3,n), gi(npg), ome(npg)
    integer :: i, n1, n2, n3, n4
    call GAULEG(-1.0d0, 1.0d0, gi, ome, npg)
    etas = NORMVEC(cone, cx, cy, cz, n, np)
    do i = 1, n
      n1 = cone(i,1); n2 = cone(i,2); n3 = cone(i,3); n4 = cone(i,4)
      cxm(i) = (cx(n1)+cx(n2)+cx(n3)+cx(n4))*0.250d0
      cym(i) = (cy(n1)+cy(n2)+cy(n3)+cy(n4))*0.250d0
      czm(i) = (cz(n1)+cz(n2)+cz(n3)+cz(n4))*0.250d0
    end do
  end subroutine

  ! The reference GHMATECE (CPU path).  consts = (pi, c1, c2, c3, c4).
  subroutine ref_ghmatece(n, nbe, np, npg, cx, cy, cz, cone, ge, rnu, hd, gd, consts) &
      bind(C, name='ref_ghmatece')
    integer(c_int), value :: n, nbe, np, npg
    real(c_double) :: cx(np), cy(np), cz(np)
    integer(c_int) :: cone(n,4)
    real(c_double), value :: ge, rnu
    real(c_double) :: hd(3,3,nbe), gd(3,3,nbe), consts(5)
    real(8), allocatable :: hest(:,:,:), gest(:,:,:)
    real(8) :: cxm(n), cym(n), czm(n), etas(3,n), gi(npg), ome(npg), delta(3,3)
    real(8) :: pi, c1, c2, c3, c4
    call set_units()
    call mesh_geometry(n, np, npg, cone, cx, cy, cz, cxm, cym, czm, etas, gi, ome)
    call GHMATECE(cx, cy, cz, cxm, cym, czm, hest, gest, n, 3*n, np, cone, n, nbe, np, &
                  npg, ge, rnu, ge, delta, pi, c1, c2, c3, c4, etas, gi, ome)
    hd = hest
    gd = gest
    consts = (/ pi, c1, c2, c3, c4 /)
  end subroutine

  ! GHMATECD's assembly loop (Ghmatecd.for:115-165) around the reference's
  ! SING_DE and NONSINGD; ZH and ZG returned before the KODE swap.
  subroutine ref_ghmatecd_raw(n, nbe, np, npg, cx, cy, cz, cone, hd, gd, fr, zge, zcs, &
                              zcp, consts, zh, zg) bind(C, name='ref_ghmatecd_raw')
    integer(c_int), value :: n, nbe, np, npg
    real(c_double) :: cx(np), cy(np), cz(np)
    integer(c_int) :: cone(n,4)
    real(c_double) :: hd(3,3,nbe), gd(3,3,nbe), consts(5)
    real(c_double), value :: fr
    complex(c_double_complex), value :: zge, zcs, zcp
    complex(c_double_complex) :: zh(3*nbe,3*nbe), zg(3*nbe,3*nbe)
    real(8) :: cxm(n), cym(n), czm(n), etas(3,n), gi(npg), ome(npg), delta(3,3), co(4,3)
    integer :: i, j, ii, jj, k
    call set_units()
    call identity3(delta)
    call mesh_geometry(n, np, npg, cone, cx, cy, cz, cxm, cym, czm, etas, gi, ome)
    !$omp parallel do private(i, j, ii, jj, k, co) schedule(dynamic, 2)
    do j = 1, nbe
      do k = 1, 4
        co(k,1) = cx(cone(j,k)); co(k,2) = cy(cone(j,k)); co(k,3) = cz(cone(j,k))
      end do
      jj = 3*(j-1) + 1
      do i = 1, nbe
        ii = 3*(i-1) + 1
        if (i == j) then
          call SING_DE(zh(ii:ii+2, jj:jj+2), zg(ii:ii+2, jj:jj+2), co, cxm(i), cym(i), &
                       czm(i), etas(1:3,j), zge, zcs, zcp, consts(2), consts(3), &
                       consts(4), consts(5), delta, consts(1), fr, gi, ome, npg)
          zg(ii:ii+2, jj:jj+2) = zg(ii:ii+2, jj:jj+2) + gd(1:3,1:3,i)
          zh(ii:ii+2, jj:jj+2) = zh(ii:ii+2, jj:jj+2) + hd(1:3,1:3,i)
        else
          call NONSINGD(zh(ii:ii+2, jj:jj+2), zg(ii:ii+2, jj:jj+2), co, cxm(i), cym(i), &
                        czm(i), etas(1:3,j), zge, zcs, zcp, delta, consts(1), fr, gi, &
                        ome, npg)
        end if
      end do
    end do
    !$omp end parallel do
  end subroutine

  ! The reference INTEREC (CPU path): un-swap, ZDSOL, ZSSOL.
  subroutine ref_interec(n, nbe, np, npg, cx, cy, cz, cone, zfi, zdfi, kode, l, cxi, cyi, &
                         czi, zge, zcs, zcp, fr, rho, pi, zdsol, zssol) &
      bind(C, name='ref_interec')
    integer(c_int), value :: n, nbe, np, npg, l
    real(c_double) :: cx(np), cy(np), cz(np)
    integer(c_int) :: cone(n,4), kode(3*nbe)
    complex(c_double_complex) :: zfi(3*nbe), zdfi(3*nbe), zdsol(3*l), zssol(9*l)
    real(c_double) :: cxi(l), cyi(l), czi(l)
    complex(c_double_complex), value :: zge, zcs, zcp
    real(c_double), value :: fr, rho, pi
    complex(8), allocatable :: dsol(:), ssol(:)
    real(8) :: cxm(n), cym(n), czm(n), etas(3,n), gi(npg), ome(npg), delta(3,3)
    call set_units()
    call identity3(delta)
    call mesh_geometry(n, np, npg, cone, cx, cy, cz, cxm, cym, czm, etas, gi, ome)
    call INTEREC(zfi, zdfi, kode, cx, cy, cz, cxi, cyi, czi, dsol, ssol, n, n, np, l, &
                 cone, zge, zcs, zcp, delta, pi, fr, npg, l, n, nbe, rho, etas, gi, ome, np)
    if (l > 0) then
      zdsol = dsol
      zssol = ssol
    end if
  end subroutine

end module bemref_harness
