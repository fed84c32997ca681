"""Python front end of libbemx.so, mirroring the reference's subroutine surface.

The four reference subroutines keep their names and argument meaning
(Ghmatece.for:9, Ghmatecd.for:11, Linsolve.for:1, Interec.for:9); arrays use
the reference's Fortran layouts (column-major, 1-based connectivity, complex
interleaved).  Every call goes to the HIP library; there is no CPU path, and
a missing or unloadable library raises at import time.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import weakref
from typing import Optional, Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libbemx.so")

MODE_FP64 = 0
MODE_MIXED = 1

STATUS = {
    0: "OK", 1: "DEGENERATE", 2: "SINGULAR", 3: "ARG", 4: "HIP", 5: "NOMEM", 6: "STATE",
    7: "NOCONV", 8: "TIMEOUT",
}

EXPORTS = [
    "bemx_create", "bemx_destroy", "bemx_last_error", "bemx_gauleg", "bemx_send_shared_data",
    "bemx_deallocate_shared_data", "bemx_ghmatece", "bemx_send_dyn_data", "bemx_ghmatecd",
    "bemx_linsolve", "bemx_interec", "bemx_solve", "bemx_solve_sweep", "bemx_device_buffers",
    "bemx_last_stats",
    "bemx_send_load_cases", "bemx_send_load_cases_kode", "bemx_material", "bemx_server_create", "bemx_server_submit",
    "bemx_server_wait", "bemx_server_destroy", "bemx_request_footprint", "bemx_server_stats",
    "bemx_linsolve_device_noalloc",
]


class BemxError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__("bemx %s (%d): %s" % (STATUS.get(code, "?"), code, msg))
        self.code = code


def load_library(path: str = LIBPATH) -> C.CDLL:
    if not os.path.exists(path):
        raise ImportError("libbemx.so is not built (run __graft_entry__.build()): %s" % path)
    lib = C.CDLL(path)
    _dp = C.POINTER(C.c_double)
    _ip = C.POINTER(C.c_int)
    vp = C.c_void_p
    lib.bemx_create.argtypes = [C.c_int, C.POINTER(vp)]
    lib.bemx_destroy.argtypes = [vp]
    lib.bemx_last_error.argtypes = [vp]
    lib.bemx_last_error.restype = C.c_char_p
    lib.bemx_gauleg.argtypes = [C.c_int, _dp, _dp]
    lib.bemx_send_shared_data.argtypes = [vp] + [_dp] * 9 + [_ip] + [C.c_int] * 4
    lib.bemx_deallocate_shared_data.argtypes = [vp]
    lib.bemx_ghmatece.argtypes = [vp, C.c_double, C.c_double] + [_dp] * 8
    lib.bemx_send_dyn_data.argtypes = [vp, _dp, _ip]
    lib.bemx_send_load_cases.argtypes = [vp, C.c_int, _dp, _ip]
    lib.bemx_send_load_cases_kode.argtypes = [vp, C.c_int, _dp, _ip]
    lib.bemx_ghmatecd.argtypes = [vp, _dp, _dp, _dp] + [C.c_double] * 5 + [_dp] * 4
    lib.bemx_linsolve.argtypes = [vp, C.c_int, _dp, _dp, C.c_int, C.POINTER(C.c_longlong)]
    lib.bemx_linsolve_device_noalloc.argtypes = [vp, C.c_int, vp, C.c_longlong, vp,
                                                 C.POINTER(C.c_longlong)]
    lib.bemx_interec.argtypes = [vp, _dp, _dp, _ip, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp,
                                 C.c_double, C.c_double, _dp, _dp]
    lib.bemx_solve.argtypes = [vp, _dp, _dp, _dp, C.c_double, C.c_int, _dp, _dp]
    lib.bemx_solve_sweep.argtypes = [vp, _dp, _dp, _dp, C.c_int, _dp, C.c_int, _dp, _dp]
    lib.bemx_device_buffers.argtypes = [vp, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp)]
    lib.bemx_last_stats.argtypes = [vp, _dp, C.c_int]
    lib.bemx_material.argtypes = [C.c_double] * 4 + [_dp] * 3
    lib.bemx_server_create.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, C.POINTER(vp)]
    lib.bemx_server_submit.argtypes = ([vp, _dp, _dp, _dp, _ip] + [C.c_int] * 4 +
                                       [C.c_double] * 5 + [C.c_int, _dp, _ip, C.c_int,
                                                           C.POINTER(C.c_longlong)])
    lib.bemx_server_wait.argtypes = [vp, C.c_longlong, _dp, _dp, C.c_char_p, C.c_int]
    lib.bemx_server_destroy.argtypes = [vp]
    lib.bemx_request_footprint.argtypes = [C.c_int] * 6
    lib.bemx_server_stats.argtypes = [vp, _dp, C.c_int]
    lib.bemx_request_footprint.restype = C.c_double
    return lib


_LIB: Optional[C.CDLL] = None

# Every live Context and Server.  They are closed at interpreter exit (before
# the interpreter's own finalisation and before the HIP runtime's exit
# handlers run), so device objects are released in a fixed order and never
# from a finaliser late in the shutdown.
_LIVE: "weakref.WeakSet" = weakref.WeakSet()


@atexit.register
def _close_all() -> None:
    for obj in list(_LIVE):
        try:
            obj.close()
        except Exception:
            pass


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        _LIB = load_library()
    return _LIB


def _p(a):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _pi(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _cz(z) -> np.ndarray:
    z = complex(z)
    return np.array([z.real, z.imag], dtype=np.float64)


def gauleg(npg: int) -> Tuple[np.ndarray, np.ndarray]:
    """GAULEG (Gauleg.for:9-36)."""
    gi = np.zeros(npg)
    ome = np.zeros(npg)
    code = lib().bemx_gauleg(npg, _p(gi), _p(ome))
    if code:
        raise BemxError(code, "gauleg")
    return gi, ome


def material(ge: float, rnu: float, dam: float, rho: float) -> np.ndarray:
    """mu* = G(1 + 2i DAM), cs = sqrt(mu*/rho), cp = cs sqrt((2-2nu)/(1-2nu))
    (Inputecd.for:58-60)."""
    zge = complex(ge, ge * 2.0 * dam)
    zcs = np.sqrt(zge / rho)
    zcp = zcs * np.sqrt((2.0 - 2.0 * rnu) / (1.0 - 2.0 * rnu))
    return np.array([zge, zcs, zcp], dtype=np.complex128)


class Context:
    """A device context holding one mesh (the reference's device globals)."""

    def __init__(self, device: int = 0):
        self._lib = lib()
        h = C.c_void_p()
        code = self._lib.bemx_create(device, C.byref(h))
        if code:
            raise BemxError(code, "bemx_create(device=%d)" % device)
        self.h = h
        self.mesh = None
        self.consts = None
        _LIVE.add(self)

    def close(self):
        h = getattr(self, "h", None)
        if h:
            self.h = None
            self._lib.bemx_destroy(h)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, code: int, what: str):
        if code:
            raise BemxError(code, "%s: %s" % (what, self._lib.bemx_last_error(self.h).decode()))

    # send_shared_data_to_gpu_ (kernels/shared.cu:84-165)
    def send_shared_data(self, mesh, cxm=None, cym=None, czm=None, etas=None, gi=None, ome=None):
        self.mesh = mesh
        cone = mesh.cone_f()
        given = etas is not None
        if given:
            etas = np.ascontiguousarray(etas, dtype=np.float64)
            cxm, cym, czm = (np.ascontiguousarray(v, dtype=np.float64) for v in (cxm, cym, czm))
        if gi is not None:
            gi = np.ascontiguousarray(gi, dtype=np.float64)
            ome = np.ascontiguousarray(ome, dtype=np.float64)
        self._keep = (cone,)
        code = self._lib.bemx_send_shared_data(
            self.h, _p(mesh.cx), _p(mesh.cy), _p(mesh.cz), _p(cxm), _p(cym), _p(czm), _p(etas),
            _p(gi), _p(ome), _pi(cone), mesh.np_, mesh.npg, mesh.n, mesh.nbe)
        self._check(code, "send_shared_data")

    def deallocate_shared_data(self):
        self._check(self._lib.bemx_deallocate_shared_data(self.h), "deallocate_shared_data")

    # GHMATECE (Ghmatece.for:9-256)
    def ghmatece(self, ge: float, rnu: float):
        """-> HESTdiag (NBE,3,3), GESTdiag (NBE,3,3), (pi, c1, c2, c3, c4)."""
        nbe = self.mesh.nbe
        hd = np.zeros(9 * nbe)
        gd = np.zeros(9 * nbe)
        delta = np.zeros(9)
        cs = [C.c_double() for _ in range(5)]
        code = self._lib.bemx_ghmatece(self.h, ge, rnu, _p(hd), _p(gd), _p(delta),
                                       *[C.byref(v) for v in cs])
        self._check(code, "ghmatece")
        self.consts = np.array([v.value for v in cs])
        return (hd.reshape(nbe, 3, 3).transpose(0, 2, 1).copy(),
                gd.reshape(nbe, 3, 3).transpose(0, 2, 1).copy(), self.consts.copy())

    # send_dyn_data_gpu_ (kernels/shared.cu:168-182)
    def send_dyn_data(self, zdfi, kode):
        zdfi = np.ascontiguousarray(zdfi, dtype=np.complex128)
        kode = np.ascontiguousarray(kode, dtype=np.int32)
        code = self._lib.bemx_send_dyn_data(self.h, _p(zdfi.view(np.float64)), _pi(kode))
        self._check(code, "send_dyn_data")
        self.nrhs = 1

    def send_load_cases(self, zdfi, kode):
        """zdfi: (3NBE, nrhs) complex, one load case per column (1..4); kode:
        (3NBE,) for every case, or (3NBE, nrhs) with a KODE column per case
        (bemx_send_load_cases_kode)."""
        kode = np.asarray(kode, dtype=np.int32)
        nn = kode.shape[0]
        zdfi = np.asfortranarray(np.asarray(zdfi, dtype=np.complex128).reshape(nn, -1))
        nrhs = zdfi.shape[1]
        zp = _p(zdfi.reshape(-1, order="F").view(np.float64))
        if kode.ndim == 2:
            kf = np.asfortranarray(kode).reshape(-1, order="F")
            code = self._lib.bemx_send_load_cases_kode(self.h, nrhs, zp, _pi(kf))
        else:
            code = self._lib.bemx_send_load_cases(self.h, nrhs, zp,
                                                  _pi(np.ascontiguousarray(kode)))
        self._check(code, "send_load_cases")
        self.nrhs = nrhs

    # GHMATECD (Ghmatecd.for:11-239)
    def ghmatecd(self, zmat, fr: float, consts=None, hestdiag=None, gestdiag=None,
                 want_zh: bool = True):
        """-> (ZH after the KODE swap or None, ZFI = ZG ZDFI)."""
        nn = 3 * self.mesh.nbe
        consts = self.consts if consts is None else consts
        zge, zcs, zcp = (_cz(v) for v in zmat)
        nr = getattr(self, "nrhs", 1)
        zfi = np.zeros(nn * nr, dtype=np.complex128)
        zh = np.zeros((nn, nn), dtype=np.complex128, order="F") if want_zh else None
        hd = gd = None
        if hestdiag is not None:
            hd = np.ascontiguousarray(np.asarray(hestdiag).transpose(0, 2, 1)).ravel()
            gd = np.ascontiguousarray(np.asarray(gestdiag).transpose(0, 2, 1)).ravel()
        code = self._lib.bemx_ghmatecd(self.h, _p(zge), _p(zcs), _p(zcp), consts[1], consts[2],
                                       consts[3], consts[4], fr, _p(hd), _p(gd),
                                       _p(zfi.view(np.float64)),
                                       _p(zh.reshape(-1, order="F").view(np.float64))
                                       if want_zh else None)
        self._check(code, "ghmatecd")
        self._last_zh = zh
        return zh, (zfi if nr == 1 else zfi.reshape(nn, nr, order="F"))

    # LINSOLVE (Linsolve.for:1-66)
    def linsolve(self, zh, zfi, mode: int = MODE_FP64):
        """Solve ZH x = ZFI; returns x.  zh=None uses the device-resident
        matrix of the last ghmatecd."""
        b = np.array(zfi, dtype=np.complex128)
        info = C.c_longlong(0)
        if zh is not None and zh is getattr(self, "_last_zh", None):
            pz = _p(zh.reshape(-1, order="F").view(np.float64))
        elif zh is not None:
            a = np.asfortranarray(zh, dtype=np.complex128)
            pz = _p(a.reshape(-1, order="F").view(np.float64))
        else:
            pz = None
        code = self._lib.bemx_linsolve(self.h, b.size, pz, _p(b.view(np.float64)), mode,
                                       C.byref(info))
        self._check(code, "linsolve (info=%d)" % info.value)
        return b

    # INTEREC (Interec.for:9-288)
    def interec(self, zfi, zdfi, kode, cxi, cyi, czi, zmat, fr: float, rho: float):
        """-> (ZFI un-swapped, ZDFI un-swapped, ZDSOL (3L), ZSSOL (9L))."""
        zfi = np.array(zfi, dtype=np.complex128)
        zdfi = np.array(zdfi, dtype=np.complex128)
        kode = np.ascontiguousarray(kode, dtype=np.int32)
        l = len(cxi)
        cxi, cyi, czi = (np.ascontiguousarray(v, dtype=np.float64) for v in (cxi, cyi, czi))
        zdsol = np.zeros(3 * l, dtype=np.complex128)
        zssol = np.zeros(9 * l, dtype=np.complex128)
        zge, zcs, zcp = (_cz(v) for v in zmat)
        code = self._lib.bemx_interec(self.h, _p(zfi.view(np.float64)), _p(zdfi.view(np.float64)),
                                      _pi(kode), l, _p(cxi), _p(cyi), _p(czi), _p(zge), _p(zcs),
                                      _p(zcp), fr, rho, _p(zdsol.view(np.float64)),
                                      _p(zssol.view(np.float64)))
        self._check(code, "interec")
        return zfi, zdfi, zdsol, zssol

    def last_stats(self) -> dict:
        out = np.zeros(11)
        self._check(self._lib.bemx_last_stats(self.h, _p(out), 11), "last_stats")
        return dict(gmres_iters=int(out[0]), relres=out[1], fell_back=bool(out[2]),
                    solve_ms=out[3], offdiag_ms=out[4], assembly_ms=out[5], lu_ms=out[6],
                    graph_captures=int(out[7]), lu_factor_ms=out[8], lu_refine_ms=out[9],
                    graph_replays=int(out[10]))

    # device-resident solve of one frequency
    def solve(self, zmat, fr: float, mode: int = MODE_FP64, want_x: bool = True):
        """-> (x, stats); x is (3NBE,) for one load case, (3NBE, nrhs) for
        several (bemx_send_load_cases)."""
        nn = 3 * self.mesh.nbe
        nr = getattr(self, "nrhs", 1)
        zge, zcs, zcp = (_cz(v) for v in zmat)
        x = np.zeros(nn * nr, dtype=np.complex128) if want_x else None
        stats = np.zeros(4)
        code = self._lib.bemx_solve(self.h, _p(zge), _p(zcs), _p(zcp), fr, mode,
                                    _p(x.view(np.float64)) if want_x else None, _p(stats))
        self._check(code, "solve")
        if want_x and nr > 1:
            x = x.reshape(nn, nr, order="F")
        return x, stats


    def solve_sweep(self, zmat, freqs, mode: int = MODE_FP64, want_x: bool = True):
        """bemx_solve_sweep: every frequency of ``freqs`` on the resident mesh
        and load cases -> (x, stats); x is (nfreq, 3NBE) for one load case,
        (nfreq, 3NBE, nrhs) for several; stats (nfreq, 4)."""
        nn = 3 * self.mesh.nbe
        nr = getattr(self, "nrhs", 1)
        f = np.ascontiguousarray(freqs, dtype=np.float64)
        zge, zcs, zcp = (_cz(v) for v in zmat)
        x = np.zeros(len(f) * nn * nr, dtype=np.complex128) if want_x else None
        stats = np.zeros(4 * len(f))
        code = self._lib.bemx_solve_sweep(self.h, _p(zge), _p(zcs), _p(zcp), len(f), _p(f), mode,
                                          _p(x.view(np.float64)) if want_x else None,
                                          _p(stats))
        self._check(code, "solve_sweep")
        if want_x:
            x = x.reshape(len(f), nr, nn).transpose(0, 2, 1) if nr > 1 else x.reshape(len(f), nn)
        return x, stats.reshape(len(f), 4)


class Server:
    """Request loop on one device (bemx_server_*): continuous batching over
    ``nslots`` worker slots with HBM-footprint admission."""

    def __init__(self, device: int = 0, nslots: int = 2, hbm_budget: float = 0.0,
                 mode: int = MODE_MIXED):
        self._lib = lib()
        self.h = None
        h = C.c_void_p()
        code = self._lib.bemx_server_create(device, nslots, float(hbm_budget), mode, C.byref(h))
        if code:
            raise BemxError(code, "bemx_server_create(nslots=%d)" % nslots)
        self.h = h
        self._sizes = {}
        _LIVE.add(self)

    def close(self):
        h = getattr(self, "h", None)
        if h:
            self.h = None
            self._lib.bemx_server_destroy(h)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, mesh, kode, zdfi, zmat_params, fr: float) -> int:
        """zdfi: (3NBE,) or (3NBE, nrhs) complex load cases; kode: (3NBE,) for
        every case or (3NBE, nrhs), a column per case; zmat_params = (ge, rnu,
        dam, rho).  Returns the request id."""
        kode = np.asarray(kode, dtype=np.int32)
        zdfi = np.asfortranarray(np.asarray(zdfi, dtype=np.complex128).reshape(kode.shape[0], -1))
        kode_cols = kode.shape[1] if kode.ndim == 2 else 1
        kode = np.ascontiguousarray(np.asfortranarray(kode).reshape(-1, order="F"))
        cx, cy, cz = (np.ascontiguousarray(v, dtype=np.float64) for v in (mesh.cx, mesh.cy, mesh.cz))
        cone = mesh.cone_f()  # column-major CONE(N,4) memory
        ge, rnu, dam, rho = (float(v) for v in zmat_params)
        rid = C.c_longlong(0)
        code = self._lib.bemx_server_submit(
            self.h, _p(cx), _p(cy), _p(cz), _pi(cone), len(cx), mesh.n, mesh.nbe, mesh.npg,
            ge, rnu, dam, rho, float(fr), zdfi.shape[1],
            _p(zdfi.reshape(-1, order="F").view(np.float64)), _pi(kode), kode_cols, C.byref(rid))
        if code:
            raise BemxError(code, "bemx_server_submit")
        self._sizes[rid.value] = zdfi.shape
        return rid.value

    def stats(self) -> dict:
        out = np.zeros(10)
        code = self._lib.bemx_server_stats(self.h, _p(out), 10)
        if code:
            raise BemxError(code, "bemx_server_stats")
        keys = ("completed", "graph_captures", "graph_replays", "charged_bytes",
                "peak_charged_bytes", "budget_bytes", "workspace_releases", "over_charge",
                "batches", "graph_precaptures")
        return {k: (int(v) if k != "budget_bytes" else float(v)) for k, v in zip(keys, out)}

    def wait(self, rid: int):
        """-> (x (3NBE,) or (3NBE, nrhs), info dict)."""
        shape = self._sizes.pop(rid)
        x = np.zeros(shape[0] * shape[1], dtype=np.complex128)
        info = np.zeros(11)
        msg = C.create_string_buffer(512)
        code = self._lib.bemx_server_wait(self.h, rid, _p(x.view(np.float64)), _p(info), msg, 512)
        if code:
            raise BemxError(code, "request %d: %s" % (rid, msg.value.decode(errors="replace")))
        if shape[1] > 1:
            x = x.reshape(shape, order="F")
        keys = ("latency_ms", "queue_ms", "service_ms", "gmres_iters", "relres", "fell_back",
                "slot", "device_ms", "start_ms", "done_ms", "batch")
        return x, dict(zip(keys, info.tolist()))


def request_footprint(mesh, nrhs: int = 1, mode: int = MODE_MIXED) -> int:
    """Bytes the request loop charges a slot that has served a request on
    ``mesh`` with ``nrhs`` load cases (bound of its context's workspace)."""
    return int(lib().bemx_request_footprint(mesh.nbe, mesh.n, mesh.np_, mesh.npg, nrhs, mode))


def material_c(ge: float, rnu: float, dam: float, rho: float) -> np.ndarray:
    """bemx_material: the library's own ZGE/ZCS/ZCP."""
    out = [np.zeros(2) for _ in range(3)]
    code = lib().bemx_material(ge, rnu, dam, rho, *[_p(o) for o in out])
    if code:
        raise BemxError(code, "bemx_material")
    return np.array([complex(o[0], o[1]) for o in out])


def solve_problem(ctx: Context, mesh, dyn, freq_index: int = 0, mode: int = MODE_FP64):
    """GHMATECE -> GHMATECD -> LINSOLVE -> INTEREC for one frequency, as
    Main.for:141-194 chains them.  Returns a dict of the results."""
    ctx.send_shared_data(mesh)
    hd, gd, consts = ctx.ghmatece(mesh.ge, mesh.rnu)
    zmat = material(mesh.ge, mesh.rnu, dyn.dam, dyn.rho)
    fr = float(dyn.freqs[freq_index])
    zdfi = dyn.bc.astype(np.complex128)
    ctx.send_dyn_data(zdfi, dyn.kode)
    zh, zfi = ctx.ghmatecd(zmat, fr)
    x = ctx.linsolve(zh, zfi, mode)
    u, t, zdsol, zssol = ctx.interec(x, zdfi, dyn.kode, dyn.cxi, dyn.cyi, dyn.czi, zmat, fr,
                                     dyn.rho)
    return dict(hestdiag=hd, gestdiag=gd, consts=consts, zmat=zmat, fr=fr, zh=zh, zfi=zfi, x=x,
                u=u, t=t, zdsol=zdsol, zssol=zssol)
