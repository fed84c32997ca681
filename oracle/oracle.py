"""ctypes front end of the test oracle.

TEST INFRASTRUCTURE ONLY -- imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker, never by the product package.

Two back ends with the same Python surface:
  * ``Oracle()``            -- liboracle.so, the C restatement (bem_oracle.c)
                               in FP64
  * ``Oracle(kind="ref")``  -- _ref/libbemref.so, the reference's own Fortran
                               routines (built here from the reference
                               sources; GHMATECD's loop restated around the
                               reference SING_DE/NONSINGD)
The dense solve (LINSOLVE, Linsolve.for:68-109 -> LAPACK ZGESV) is numpy's
LAPACK zgesv for both.

``OracleLD`` is the same C source built in x87 long double (64-bit
significand): GHMATECE, GHMATECD (full or sampled blocks) and SOLFUND on the
same double inputs, every operation in extended precision, results rounded to
double.  It is the reference value where the FP64 formulas of the dynamic
kernels lose digits (k R << 1: their 1/(kR)^2 terms cancel between the S and
P waves, so FP64 rounding grows like eps / (kR)^2).

Array conventions are the reference's (column-major, 1-based connectivity):
3x3 blocks come back as (3,3) arrays indexed [i, j]; HESTdiag/GESTdiag as
(NBE, 3, 3); ZH as an (n, n) complex128 Fortran-ordered array.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {"c": os.path.join(HERE, "liboracle.so"),
         "ld": os.path.join(HERE, "liboracle_ld.so"),
         "ref": os.path.join(HERE, "_ref", "libbemref.so"),
         "ref_fast": os.path.join(HERE, "_ref", "libbemref_fast.so")}

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class _Z(C.Structure):
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


def _z(v) -> _Z:
    v = complex(v)
    return _Z(v.real, v.imag)


def _p(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _pi(a: np.ndarray):
    return a.ctypes.data_as(_ip)


def build(ref: bool = False) -> None:
    """Compile the oracle (and, with ref=True, the reference build and the
    reference's result checker, dtester)."""
    targets = ["all"] + (["ref", "dtester"] if ref else [])
    subprocess.run(["make", "-s", "-C", HERE] + targets, check=True)


def ref_available() -> bool:
    return os.path.exists(_LIBS["ref"])


def _fmat(a, shape=(3, 3)):
    return np.asarray(a).reshape(shape[::-1]).T.copy()


class Oracle:
    def __init__(self, kind: str = "c"):
        path = _LIBS[kind]
        if not os.path.exists(path):
            raise FileNotFoundError("%s not built (make -C oracle%s)" %
                                    (path, " ref" if kind == "ref" else ""))
        self.kind = "ref" if kind == "ref_fast" else kind
        self.lib = C.CDLL(path)

    # ---------------- setup -----------------
    def gauleg(self, npg: int):
        gi = np.zeros(npg)
        ome = np.zeros(npg)
        if self.kind == "c":
            self.lib.orc_gauleg(C.c_int(npg), _p(gi), _p(ome))
        else:
            self.lib.ref_gauleg(C.c_int(npg), _p(gi), _p(ome))
        return gi, ome

    def normvec(self, mesh):
        cone = mesh.cone_f()
        et = np.zeros(3 * mesh.n)
        f = self.lib.orc_normvec if self.kind == "c" else self.lib.ref_normvec
        f(C.c_int(mesh.n), C.c_int(mesh.np_), _pi(cone), _p(mesh.cx), _p(mesh.cy),
          _p(mesh.cz), _p(et))
        return et.reshape(mesh.n, 3)

    def material(self, ge, rnu, dam, rho):
        out = np.zeros(3, dtype=np.complex128)
        f = self.lib.orc_material if self.kind == "c" else self.lib.ref_material
        f.argtypes = [C.c_double] * 4 + [_dp]
        f(ge, rnu, dam, rho, _p(out))
        return out

    # ---------------- point kernels -----------------
    def solfund(self, xp, xg, rn, zge, zcs, zcp, pi, fr):
        xp, xg, rn = (np.ascontiguousarray(v, dtype=np.float64) for v in (xp, xg, rn))
        zu = np.zeros(9, dtype=np.complex128)
        zt = np.zeros(9, dtype=np.complex128)
        f = self.lib.orc_solfund if self.kind == "c" else self.lib.ref_solfund
        f.argtypes = [_dp, _dp, _dp, _Z, _Z, _Z, C.c_double, C.c_double, _dp, _dp]
        f(_p(xp), _p(xg), _p(rn), _z(zge), _z(zcs), _z(zcp), pi, fr, _p(zu), _p(zt))
        return _fmat(zu), _fmat(zt)

    def solfundif(self, xp, xg, rn, zge, zcs, zcp, c1, c2, c3, c4, pi, fr):
        xp, xg, rn = (np.ascontiguousarray(v, dtype=np.float64) for v in (xp, xg, rn))
        zu = np.zeros(9, dtype=np.complex128)
        zt = np.zeros(9, dtype=np.complex128)
        if self.kind == "c":
            f = self.lib.orc_solfundif
            f.argtypes = [_dp, _dp, _dp, _Z, _Z, _Z] + [C.c_double] * 6 + [_dp, _dp]
            f(_p(xp), _p(xg), _p(rn), _z(zge), _z(zcs), _z(zcp), c1, c2, c3, c4, pi, fr,
              _p(zu), _p(zt))
        else:
            zm = np.array([zge, zcs, zcp], dtype=np.complex128)
            f = self.lib.ref_solfundif
            f.argtypes = [_dp, _dp, _dp, _dp] + [C.c_double] * 6 + [_dp, _dp]
            f(_p(xp), _p(xg), _p(rn), _p(zm), c1, c2, c3, c4, pi, fr, _p(zu), _p(zt))
        return _fmat(zu), _fmat(zt)

    def solfune(self, xp, xg, eta, c3, c4):
        xp, xg, eta = (np.ascontiguousarray(v, dtype=np.float64) for v in (xp, xg, eta))
        t = np.zeros(9)
        f = self.lib.orc_solfune if self.kind == "c" else self.lib.ref_solfune
        f.argtypes = [_dp, _dp, _dp, C.c_double, C.c_double, _dp]
        f(_p(xp), _p(xg), _p(eta), c3, c4, _p(t))
        return _fmat(t)

    # ---------------- element quadratures -----------------
    @staticmethod
    def _co(co):
        """CO(4,3) column-major from a (4,3) corner array."""
        return np.asfortranarray(np.asarray(co, dtype=np.float64)).ravel(order="F")

    def nonsinge(self, co, xp, eta, c3, c4, gi, ome):
        h = np.zeros(9)
        xp, eta = (np.ascontiguousarray(v, dtype=np.float64) for v in (xp, eta))
        npg = len(gi)
        gi, ome = np.ascontiguousarray(gi), np.ascontiguousarray(ome)
        if self.kind == "c":
            co_c = np.ascontiguousarray(co, dtype=np.float64)
            f = self.lib.orc_nonsinge
            f.argtypes = [_dp, _dp, _dp, C.c_double, C.c_double, _dp, _dp, C.c_int, _dp]
            st = f(_p(co_c), _p(xp), _p(eta), c3, c4, _p(gi), _p(ome), npg, _p(h))
            if st:
                raise ArithmeticError("degenerate Jacobian")
        else:
            co_f = self._co(co)
            f = self.lib.ref_nonsinge
            f.argtypes = [_dp, _dp, _dp, C.c_double, C.c_double, C.c_int, _dp, _dp, _dp]
            f(_p(co_f), _p(xp), _p(eta), c3, c4, npg, _p(gi), _p(ome), _p(h))
        return _fmat(h)

    def singge(self, co, xm, c1, c2, gi, ome):
        g = np.zeros(9)
        xm = np.ascontiguousarray(xm, dtype=np.float64)
        npg = len(gi)
        gi, ome = np.ascontiguousarray(gi), np.ascontiguousarray(ome)
        if self.kind == "c":
            co_c = np.ascontiguousarray(co, dtype=np.float64)
            f = self.lib.orc_singge
            f.argtypes = [_dp, _dp, C.c_double, C.c_double, _dp, _dp, C.c_int, _dp]
            if f(_p(co_c), _p(xm), c1, c2, _p(gi), _p(ome), npg, _p(g)):
                raise ArithmeticError("degenerate Jacobian")
        else:
            co_f = self._co(co)
            f = self.lib.ref_singge
            f.argtypes = [_dp, _dp, C.c_double, C.c_double, C.c_int, _dp, _dp, _dp]
            f(_p(co_f), _p(xm), c1, c2, npg, _p(gi), _p(ome), _p(g))
        return _fmat(g)

    def nonsingd(self, co, xp, eta, zge, zcs, zcp, pi, fr, gi, ome):
        zh = np.zeros(9, dtype=np.complex128)
        zg = np.zeros(9, dtype=np.complex128)
        xp, eta = (np.ascontiguousarray(v, dtype=np.float64) for v in (xp, eta))
        npg = len(gi)
        gi, ome = np.ascontiguousarray(gi), np.ascontiguousarray(ome)
        if self.kind == "c":
            co_c = np.ascontiguousarray(co, dtype=np.float64)
            f = self.lib.orc_nonsingd
            f.argtypes = [_dp, _dp, _dp, _Z, _Z, _Z, C.c_double, C.c_double, _dp, _dp,
                          C.c_int, _dp, _dp]
            if f(_p(co_c), _p(xp), _p(eta), _z(zge), _z(zcs), _z(zcp), pi, fr, _p(gi),
                 _p(ome), npg, _p(zh), _p(zg)):
                raise ArithmeticError("degenerate Jacobian")
        else:
            co_f = self._co(co)
            f = self.lib.ref_nonsingd
            f.argtypes = [_dp, _dp, _dp, _Z, _Z, _Z, C.c_double, C.c_double, C.c_int,
                          _dp, _dp, _dp, _dp]
            f(_p(co_f), _p(xp), _p(eta), _z(zge), _z(zcs), _z(zcp), pi, fr, npg, _p(gi),
              _p(ome), _p(zh), _p(zg))
        return _fmat(zh), _fmat(zg)

    def sigmaec(self, co, xp, rn, pi, fr, zge, rho, zcs, zcp, gi, ome):
        """Returns ZD, ZS as (3,3,3) arrays indexed [k, i, j]."""
        zd = np.zeros(27, dtype=np.complex128)
        zs = np.zeros(27, dtype=np.complex128)
        xp, rn = (np.ascontiguousarray(v, dtype=np.float64) for v in (xp, rn))
        npg = len(gi)
        gi, ome = np.ascontiguousarray(gi), np.ascontiguousarray(ome)
        if self.kind == "c":
            co_c = np.ascontiguousarray(co, dtype=np.float64)
            f = self.lib.orc_sigmaec
            f.argtypes = [_dp, _dp, _dp, C.c_double, C.c_double, _Z, C.c_double, _Z, _Z,
                          _dp, _dp, C.c_int, _dp, _dp]
            if f(_p(co_c), _p(xp), _p(rn), pi, fr, _z(zge), rho, _z(zcs), _z(zcp), _p(gi),
                 _p(ome), npg, _p(zd), _p(zs)):
                raise ArithmeticError("degenerate Jacobian")
        else:
            co_f = self._co(co)
            zm = np.array([zge, zcs, zcp], dtype=np.complex128)
            f = self.lib.ref_sigmaec
            f.argtypes = [_dp, _dp, _dp, C.c_double, C.c_double, _dp, C.c_double, C.c_int,
                          _dp, _dp, _dp, _dp]
            f(_p(co_f), _p(xp), _p(rn), pi, fr, _p(zm), rho, npg, _p(gi), _p(ome), _p(zd),
              _p(zs))
        return (zd.reshape(3, 3, 3).transpose(2, 1, 0).copy(),
                zs.reshape(3, 3, 3).transpose(2, 1, 0).copy())

    # ---------------- assembly -----------------
    def ghmatece(self, mesh):
        """-> hestdiag (NBE,3,3), gestdiag (NBE,3,3), consts (pi,c1,c2,c3,c4)."""
        nbe = mesh.nbe
        hd = np.zeros(9 * nbe)
        gd = np.zeros(9 * nbe)
        consts = np.zeros(5)
        cone = mesh.cone_f()
        if self.kind == "c":
            f = self.lib.orc_ghmatece
            f.argtypes = [C.c_int] * 4 + [_dp, _dp, _dp, _ip, C.c_double, C.c_double,
                                          _dp, _dp, _dp, _dp]
            st = f(mesh.n, nbe, mesh.np_, mesh.npg, _p(mesh.cx), _p(mesh.cy), _p(mesh.cz),
                   _pi(cone), mesh.ge, mesh.rnu, _p(hd), _p(gd), _p(consts), None)
            if st:
                raise ArithmeticError("GHMATECE status %d" % st)
        else:
            f = self.lib.ref_ghmatece
            f.argtypes = [C.c_int] * 4 + [_dp, _dp, _dp, _ip, C.c_double, C.c_double,
                                          _dp, _dp, _dp]
            f(mesh.n, nbe, mesh.np_, mesh.npg, _p(mesh.cx), _p(mesh.cy), _p(mesh.cz),
              _pi(cone), mesh.ge, mesh.rnu, _p(hd), _p(gd), _p(consts))
        return (hd.reshape(nbe, 3, 3).transpose(0, 2, 1).copy(),
                gd.reshape(nbe, 3, 3).transpose(0, 2, 1).copy(), consts)

    @staticmethod
    def _diag_raw(d):
        return np.ascontiguousarray(np.asarray(d).transpose(0, 2, 1)).ravel()

    def ghmatecd(self, mesh, hestdiag, gestdiag, consts, kode, zdfi, fr, zmat,
                 return_zg=False):
        """GHMATECD -> (ZH after the KODE swap, ZFI[, ZG after the swap])."""
        nbe = mesh.nbe
        nn = 3 * nbe
        zge, zcs, zcp = (complex(v) for v in zmat)
        hd, gd = self._diag_raw(hestdiag), self._diag_raw(gestdiag)
        zh = np.zeros((nn, nn), dtype=np.complex128, order="F")
        zg = np.zeros((nn, nn), dtype=np.complex128, order="F")
        kode = np.ascontiguousarray(kode, dtype=np.int32)
        zdfi = np.ascontiguousarray(zdfi, dtype=np.complex128)
        cone = mesh.cone_f()
        consts = np.ascontiguousarray(consts, dtype=np.float64)
        if self.kind == "c":
            zfi = np.zeros(nn, dtype=np.complex128)
            f = self.lib.orc_ghmatecd
            f.argtypes = [C.c_int] * 4 + [_dp, _dp, _dp, _ip, _dp, _dp, _ip, _dp,
                                          C.c_double, _Z, _Z, _Z, _dp, _dp, _dp, _dp]
            st = f(mesh.n, nbe, mesh.np_, mesh.npg, _p(mesh.cx), _p(mesh.cy), _p(mesh.cz),
                   _pi(cone), _p(hd), _p(gd), _pi(kode), _p(zdfi), fr, _z(zge), _z(zcs),
                   _z(zcp), _p(consts), _p(zh), _p(zfi), _p(zg))
            if st:
                raise ArithmeticError("GHMATECD status %d" % st)
        else:
            f = self.lib.ref_ghmatecd_raw
            f.argtypes = [C.c_int] * 4 + [_dp, _dp, _dp, _ip, _dp, _dp, C.c_double, _Z,
                                          _Z, _Z, _dp, _dp, _dp]
            f(mesh.n, nbe, mesh.np_, mesh.npg, _p(mesh.cx), _p(mesh.cy), _p(mesh.cz),
              _pi(cone), _p(hd), _p(gd), fr, _z(zge), _z(zcs), _z(zcp), _p(consts),
              _p(zh), _p(zg))
            # BOUNDARY_CONDITIONS (Ghmatecd.for:168,316-333) and ZGEMV (:172-174)
            for c in np.nonzero(kode == 0)[0]:
                tmp = zh[:, c] * zge
                zh[:, c] = -zg[:, c]
                zg[:, c] = -tmp
            zfi = zg @ zdfi
        return (zh, zfi, zg) if return_zg else (zh, zfi)

    def ghmatecd_blocks(self, mesh, hestdiag, gestdiag, consts, fr, zmat, bi, bj):
        """3x3 blocks (bi[k], bj[k]) of ZH and ZG before the KODE swap, as the
        full GHMATECD forms them -> (H, G), each (k, 3, 3) indexed [k, a, b]
        (C oracle only)."""
        if self.kind != "c":
            raise NotImplementedError("block sampling is in the C oracle")
        bi = np.ascontiguousarray(bi, dtype=np.int32)
        bj = np.ascontiguousarray(bj, dtype=np.int32)
        k = len(bi)
        hb = np.zeros(9 * k, dtype=np.complex128)
        gb = np.zeros(9 * k, dtype=np.complex128)
        zge, zcs, zcp = (complex(v) for v in zmat)
        hd, gd = self._diag_raw(hestdiag), self._diag_raw(gestdiag)
        consts = np.ascontiguousarray(consts, dtype=np.float64)
        cone = mesh.cone_f()
        f = self.lib.orc_ghmatecd_blocks
        f.argtypes = [C.c_int] * 4 + [_dp, _dp, _dp, _ip, _dp, _dp, C.c_double, _Z, _Z, _Z,
                                      _dp, C.c_int, _ip, _ip, _dp, _dp]
        st = f(mesh.n, mesh.nbe, mesh.np_, mesh.npg, _p(mesh.cx), _p(mesh.cy), _p(mesh.cz),
               _pi(cone), _p(hd), _p(gd), fr, _z(zge), _z(zcs), _z(zcp), _p(consts), k,
               _pi(bi), _pi(bj), _p(hb), _p(gb))
        if st:
            raise ArithmeticError("GHMATECD blocks status %d" % st)
        return (hb.reshape(k, 3, 3).transpose(0, 2, 1).copy(),
                gb.reshape(k, 3, 3).transpose(0, 2, 1).copy())

    def ghmatecd_accurate(self, mesh, hestdiag, gestdiag, consts, kode, zdfi, fr, zmat,
                          full_ld=False):
        """GHMATECD with the entries the FP64 formulas cannot give to 1e-12
        taken from the extended-precision build: the diagonal blocks (the
        dynamic-minus-static difference kernels, Sing_de.for) always, every
        block with full_ld.  Off the diagonal the FP64 kernels are accurate
        to ~1e-13 while k R >~ 0.3 over the near field (every benchmark mesh
        from omega ~ 1 up).  -> (ZH, ZFI, ZG) after the KODE swap, ZFI
        recomputed from ZG (C oracle only)."""
        ld = OracleLD()
        nbe, nn = mesh.nbe, 3 * mesh.nbe
        zdfi = np.ascontiguousarray(zdfi, dtype=np.complex128)
        kode = np.ascontiguousarray(kode, dtype=np.int32)
        zge = complex(zmat[0])
        if full_ld:
            # the extended-precision GHMATECD swaps and forms ZFI itself
            zh, zfi = ld.ghmatecd(mesh, hestdiag, gestdiag, consts, kode, zdfi, fr, zmat)
            return zh, zfi, None
        else:
            ones = np.ones(nn, dtype=np.int32)
            zh, _, zg = self.ghmatecd(mesh, hestdiag, gestdiag, consts, ones, zdfi, fr, zmat,
                                      return_zg=True)
            d = np.arange(nbe)
            hb, gb = ld.ghmatecd_blocks(mesh, hestdiag, gestdiag, consts, fr, zmat, d, d)
            for i in range(nbe):
                zh[3 * i:3 * i + 3, 3 * i:3 * i + 3] = hb[i]
                zg[3 * i:3 * i + 3, 3 * i:3 * i + 3] = gb[i]
        # BOUNDARY_CONDITIONS (Ghmatecd.for:168,316-333) and ZFI = ZG ZDFI (:172-174)
        for c in np.nonzero(kode == 0)[0]:
            tmp = zh[:, c] * zge
            zh[:, c] = -zg[:, c]
            zg[:, c] = -tmp
        return zh, zg @ zdfi, zg

    @staticmethod
    def linsolve(zh, zfi):
        """LINSOLVE: LAPACK zgesv through numpy (Linsolve.for:68-109)."""
        return np.linalg.solve(zh, zfi)

    def zgesv_plain(self, zh, zfi):
        """In-oracle partial-pivot LU (small systems only)."""
        a = np.array(zh, dtype=np.complex128, order="F")
        b = np.array(zfi, dtype=np.complex128)
        info = self.lib.orc_zgesv(C.c_int(a.shape[0]), _p(a), _p(b))
        if info:
            raise np.linalg.LinAlgError("singular at column %d" % info)
        return b

    def interec(self, mesh, dyn, zfi, zdfi, fr, zmat, pi):
        """INTEREC -> (ZFI un-swapped, ZDFI un-swapped, ZDSOL (3L), ZSSOL (9L))."""
        nbe, l = mesh.nbe, dyn.l
        zfi = np.array(zfi, dtype=np.complex128)
        zdfi = np.array(zdfi, dtype=np.complex128)
        zdsol = np.zeros(3 * l, dtype=np.complex128)
        zssol = np.zeros(9 * l, dtype=np.complex128)
        cone = mesh.cone_f()
        kode = np.ascontiguousarray(dyn.kode, dtype=np.int32)
        zge, zcs, zcp = (complex(v) for v in zmat)
        cxi, cyi, czi = (np.ascontiguousarray(v, dtype=np.float64)
                         for v in (dyn.cxi, dyn.cyi, dyn.czi))
        if self.kind == "c":
            f = self.lib.orc_interec
            f.argtypes = [C.c_int] * 4 + [_dp, _dp, _dp, _ip, _dp, _dp, _ip, C.c_int, _dp,
                                          _dp, _dp, _Z, _Z, _Z, C.c_double, C.c_double,
                                          C.c_double, _dp, _dp]
            st = f(mesh.n, nbe, mesh.np_, mesh.npg, _p(mesh.cx), _p(mesh.cy), _p(mesh.cz),
                   _pi(cone), _p(zfi), _p(zdfi), _pi(kode), l, _p(cxi), _p(cyi), _p(czi),
                   _z(zge), _z(zcs), _z(zcp), fr, dyn.rho, pi, _p(zdsol), _p(zssol))
            if st:
                raise ArithmeticError("INTEREC status %d" % st)
        else:
            f = self.lib.ref_interec
            f.argtypes = [C.c_int] * 4 + [_dp, _dp, _dp, _ip, _dp, _dp, _ip, C.c_int, _dp,
                                          _dp, _dp, _Z, _Z, _Z, C.c_double, C.c_double,
                                          C.c_double, _dp, _dp]
            f(mesh.n, nbe, mesh.np_, mesh.npg, _p(mesh.cx), _p(mesh.cy), _p(mesh.cz),
              _pi(cone), _p(zfi), _p(zdfi), _pi(kode), l, _p(cxi), _p(cyi), _p(czi),
              _z(zge), _z(zcs), _z(zcp), fr, dyn.rho, pi, _p(zdsol), _p(zssol))
        return zfi, zdfi, zdsol, zssol

    # ---------------- whole chain -----------------
    def solve_problem(self, mesh, dyn, freq_index=0, zdfi=None, accurate=False,
                      full_ld=False):
        """GHMATECE -> GHMATECD -> LINSOLVE -> INTEREC for one frequency.
        Returns a dict of every intermediate (Main.for:141-194).  accurate:
        GHMATECD through ghmatecd_accurate (diagonal blocks, or with full_ld
        every block, from the extended-precision build)."""
        hd, gd, consts = self.ghmatece(mesh)
        zmat = self.material(mesh.ge, mesh.rnu, dyn.dam, dyn.rho)
        fr = float(dyn.freqs[freq_index])
        if zdfi is None:
            zdfi = dyn.bc.astype(np.complex128)
        if accurate:
            zh, zfi, zg = self.ghmatecd_accurate(mesh, hd, gd, consts, dyn.kode, zdfi, fr, zmat,
                                                 full_ld=full_ld)
        else:
            zh, zfi, zg = self.ghmatecd(mesh, hd, gd, consts, dyn.kode, zdfi, fr, zmat,
                                        return_zg=True)
        x = self.linsolve(zh, zfi)
        u, t, zdsol, zssol = self.interec(mesh, dyn, x, zdfi, fr, zmat, consts[0])
        return dict(hestdiag=hd, gestdiag=gd, consts=consts, zmat=zmat, fr=fr, zh=zh,
                    zg=zg, zfi=zfi, x=x, u=u, t=t, zdsol=zdsol, zssol=zssol)


def quad_rule(orc: "Oracle", co, gi, ome):
    """Gauss points (NPG^2, 3) and weights OME*OME*DET of one element."""
    npg = len(gi)
    xg = np.zeros(3 * npg * npg)
    w = np.zeros(npg * npg)
    co_c = np.ascontiguousarray(co, dtype=np.float64)
    gi, ome = np.ascontiguousarray(gi), np.ascontiguousarray(ome)
    f = orc.lib.orc_quad_rule
    f.argtypes = [_dp, _dp, _dp, C.c_int, _dp, _dp]
    if f(_p(co_c), _p(gi), _p(ome), npg, _p(xg), _p(w)):
        raise ArithmeticError("degenerate Jacobian")
    return xg.reshape(-1, 3), w


class OracleLD:
    """The C restatement evaluated in x87 extended precision (liboracle_ld.so);
    same arguments and array conventions as Oracle("c")."""

    def __init__(self):
        path = _LIBS["ld"]
        if not os.path.exists(path):
            raise FileNotFoundError("%s not built (make -C oracle)" % path)
        self.lib = C.CDLL(path)

    @staticmethod
    def _zmat(zmat):
        return np.array([v for z in zmat for v in (complex(z).real, complex(z).imag)])

    def ghmatece(self, mesh):
        nbe = mesh.nbe
        hd, gd, consts = np.zeros(9 * nbe), np.zeros(9 * nbe), np.zeros(5)
        f = self.lib.orcld_ghmatece
        f.argtypes = [C.c_int] * 4 + [_dp, _dp, _dp, _ip, C.c_double, C.c_double, _dp, _dp, _dp]
        if f(mesh.n, nbe, mesh.np_, mesh.npg, _p(mesh.cx), _p(mesh.cy), _p(mesh.cz),
             _pi(mesh.cone_f()), mesh.ge, mesh.rnu, _p(hd), _p(gd), _p(consts)):
            raise ArithmeticError("GHMATECE (long double)")
        return (hd.reshape(nbe, 3, 3).transpose(0, 2, 1).copy(),
                gd.reshape(nbe, 3, 3).transpose(0, 2, 1).copy(), consts)

    def ghmatecd(self, mesh, hestdiag, gestdiag, consts, kode, zdfi, fr, zmat):
        """-> (ZH after the KODE swap, ZFI = ZG ZDFI)."""
        nn = 3 * mesh.nbe
        zh = np.zeros((nn, nn), dtype=np.complex128, order="F")
        zfi = np.zeros(nn, dtype=np.complex128)
        zdfi = np.ascontiguousarray(zdfi, dtype=np.complex128)
        f = self.lib.orcld_ghmatecd
        f.argtypes = [C.c_int] * 4 + [_dp, _dp, _dp, _ip, _dp, _dp, _ip, _dp, C.c_double, _dp,
                                      _dp, _dp, _dp]
        if f(mesh.n, mesh.nbe, mesh.np_, mesh.npg, _p(mesh.cx), _p(mesh.cy), _p(mesh.cz),
             _pi(mesh.cone_f()), _p(Oracle._diag_raw(hestdiag)), _p(Oracle._diag_raw(gestdiag)),
             _pi(np.ascontiguousarray(kode, dtype=np.int32)), _p(zdfi.view(np.float64)), fr,
             _p(self._zmat(zmat)), _p(np.ascontiguousarray(consts, dtype=np.float64)),
             _p(zh.reshape(-1, order="F").view(np.float64)), _p(zfi.view(np.float64))):
            raise ArithmeticError("GHMATECD (long double)")
        return zh, zfi

    def ghmatecd_blocks(self, mesh, hestdiag, gestdiag, consts, fr, zmat, bi, bj):
        """-> (H, G) blocks (k, 3, 3) before the KODE swap, as Oracle.ghmatecd_blocks."""
        bi = np.ascontiguousarray(bi, dtype=np.int32)
        bj = np.ascontiguousarray(bj, dtype=np.int32)
        k = len(bi)
        hb = np.zeros(9 * k, dtype=np.complex128)
        gb = np.zeros(9 * k, dtype=np.complex128)
        f = self.lib.orcld_ghmatecd_blocks
        f.argtypes = [C.c_int] * 4 + [_dp, _dp, _dp, _ip, _dp, _dp, C.c_double, _dp, _dp,
                                      C.c_int, _ip, _ip, _dp, _dp]
        if f(mesh.n, mesh.nbe, mesh.np_, mesh.npg, _p(mesh.cx), _p(mesh.cy), _p(mesh.cz),
             _pi(mesh.cone_f()), _p(Oracle._diag_raw(hestdiag)), _p(Oracle._diag_raw(gestdiag)),
             fr, _p(self._zmat(zmat)), _p(np.ascontiguousarray(consts, dtype=np.float64)), k,
             _pi(bi), _pi(bj), _p(hb.view(np.float64)), _p(gb.view(np.float64))):
            raise ArithmeticError("GHMATECD blocks (long double)")
        return (hb.reshape(k, 3, 3).transpose(0, 2, 1).copy(),
                gb.reshape(k, 3, 3).transpose(0, 2, 1).copy())

    def solfund(self, xp, xg, rn, zge, zcs, zcp, pi, fr):
        xp, xg, rn = (np.ascontiguousarray(v, dtype=np.float64) for v in (xp, xg, rn))
        zu = np.zeros(9, dtype=np.complex128)
        zt = np.zeros(9, dtype=np.complex128)
        f = self.lib.orcld_solfund
        f.argtypes = [_dp, _dp, _dp, _dp, C.c_double, C.c_double, _dp, _dp]
        f(_p(xp), _p(xg), _p(rn), _p(self._zmat((zge, zcs, zcp))), pi, fr,
          _p(zu.view(np.float64)), _p(zt.view(np.float64)))
        return _fmat(zu), _fmat(zt)
