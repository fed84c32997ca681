"""Parity of the HIP assembly and interior kernels with the oracle (GPU).

Bar: FP64 assembly agrees per entry to 1e-12 relative, with the norm floor
of conftest.close_blocks (an entry is judged against max(|entry|, max of its
3x3 block)); vectors to 1e-12 relative to their max entry.  The dynamic
matrices are the oracle's with its diagonal blocks from the extended-
precision build (Oracle.ghmatecd_accurate): the FP64 difference kernels of
those blocks lose digits at small k R (eps / (kR)^2) beyond this bar.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, close_blocks
from bemx import api, decks

pytestmark = pytest.mark.gpu

RTOL = 1e-12


@pytest.fixture(scope="module")
def ctx():
    c = api.Context(0)
    yield c
    c.close()


def _problem(case):
    if case == "c1":
        return decks.config_problem("c1")
    if case == "esolo240":
        mesh = decks.read_e_deck(os.path.join(GOLDEN, "ESOLO240E_-5+5.DAT"))
        dyn = decks.read_d_deck(os.path.join(GOLDEN, "ESOLO240D_-5+5.DAT"), mesh.nbe)
        return mesh, dyn
    if case == "npg6":
        return decks.flat_patch_problem(12, 11, half=4.0, npg=6, ge=1.7, rnu=0.31, dam=0.03,
                                        freqs=(3.5,), n_interior=9, seed=3)
    if case == "hifreq":
        # about 40 shear wavelengths across the patch: distances up to 22.6 at
        # |ks| = 12 (the H kernel's exp table then holds ~5800 entries per wave)
        return decks.flat_patch_problem(16, 15, half=8.0, dam=0.05, freqs=(12.0,),
                                        n_interior=4, seed=5)
    raise KeyError(case)


def _vec_close(a, b, rtol=RTOL):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / np.max(np.abs(b)))


@pytest.mark.parametrize("case", ["c1", "esolo240", "npg6", "hifreq"])
def test_static_and_dynamic_assembly(ctx, orc, case):
    mesh, dyn = _problem(case)
    print("case", case, "nbe", mesh.nbe, "n", mesh.n, "npg", mesh.npg, flush=True)
    ctx.send_shared_data(mesh)
    hd, gd, consts = ctx.ghmatece(mesh.ge, mesh.rnu)
    ohd, ogd, oconsts = orc.ghmatece(mesh)
    np.testing.assert_allclose(consts, oconsts, rtol=1e-15)
    ok, worst = close_blocks(hd.reshape(-1, 9), ohd.reshape(-1, 9), RTOL)
    assert ok, ("HESTdiag", worst)
    ok, worst = close_blocks(gd.reshape(-1, 9), ogd.reshape(-1, 9), RTOL)
    assert ok, ("GESTdiag", worst)

    zmat = api.material(mesh.ge, mesh.rnu, dyn.dam, dyn.rho)
    fr = float(dyn.freqs[0])
    zdfi = dyn.bc.astype(np.complex128)
    ctx.send_dyn_data(zdfi, dyn.kode)
    zh, zfi = ctx.ghmatecd(zmat, fr)
    ozh, ozfi, _ = orc.ghmatecd_accurate(mesh, ohd, ogd, oconsts, dyn.kode, zdfi, fr, zmat)
    ok, worst = close_blocks(zh, ozh, RTOL)
    assert ok, ("ZH", worst)
    assert _vec_close(zfi, ozfi) < RTOL


@pytest.mark.parametrize("case", ["c1", "hifreq"])
def test_dynamic_assembly_without_exp_table(ctx, orc, case, monkeypatch):
    # BEMX_EXP_TABLE=0: the H kernel evaluates exp and sincos at every Gauss
    # point, the route it takes when the table would exceed its maximum size
    monkeypatch.setenv("BEMX_EXP_TABLE", "0")
    mesh, dyn = _problem(case)
    print("case", case, "nbe", mesh.nbe, "without exp table", flush=True)
    ctx.send_shared_data(mesh)
    hd, gd, consts = ctx.ghmatece(mesh.ge, mesh.rnu)
    ohd, ogd, oconsts = orc.ghmatece(mesh)
    zmat = api.material(mesh.ge, mesh.rnu, dyn.dam, dyn.rho)
    fr = float(dyn.freqs[0])
    zdfi = dyn.bc.astype(np.complex128)
    ctx.send_dyn_data(zdfi, dyn.kode)
    zh, zfi = ctx.ghmatecd(zmat, fr)
    ozh, ozfi, _ = orc.ghmatecd_accurate(mesh, ohd, ogd, oconsts, dyn.kode, zdfi, fr, zmat)
    ok, worst = close_blocks(zh, ozh, RTOL)
    assert ok, ("ZH", worst)
    assert _vec_close(zfi, ozfi) < RTOL


@pytest.mark.parametrize("case", ["c1", "esolo240", "npg6"])
def test_interior_points(ctx, orc, case):
    mesh, dyn = _problem(case)
    o = orc.solve_problem(mesh, dyn)
    ctx.send_shared_data(mesh)
    ctx.ghmatece(mesh.ge, mesh.rnu)
    zdfi = dyn.bc.astype(np.complex128)
    u, t, zdsol, zssol = ctx.interec(o["x"], zdfi, dyn.kode, dyn.cxi, dyn.cyi, dyn.czi,
                                     o["zmat"], o["fr"], dyn.rho)
    assert _vec_close(zdsol, o["zdsol"]) < RTOL
    assert _vec_close(zssol, o["zssol"]) < RTOL
    np.testing.assert_array_equal(u, o["u"])
    np.testing.assert_array_equal(t, o["t"])


def test_kode_zero_columns(ctx, orc):
    """Displacement-prescribed DOFs: the mu*-scaled column swap and the RHS."""
    mesh, dyn = decks.flat_patch_problem(8, 7, n_interior=5, freqs=(1.7,))
    rng = np.random.default_rng(11)
    kode = dyn.kode.copy()
    sel = rng.choice(kode.size, 20, replace=False)
    kode[sel] = 0
    zdfi = dyn.bc.astype(np.complex128)
    zdfi[sel] = rng.normal(size=20) + 1j * rng.normal(size=20)
    ctx.send_shared_data(mesh)
    hd, gd, consts = ctx.ghmatece(mesh.ge, mesh.rnu)
    zmat = api.material(mesh.ge, mesh.rnu, dyn.dam, dyn.rho)
    ctx.send_dyn_data(zdfi, kode)
    zh, zfi = ctx.ghmatecd(zmat, 1.7)
    ohd, ogd, oconsts = orc.ghmatece(mesh)
    ozh, ozfi, _ = orc.ghmatecd_accurate(mesh, ohd, ogd, oconsts, kode, zdfi, 1.7, zmat)
    ok, worst = close_blocks(zh, ozh, RTOL)
    assert ok, worst
    assert _vec_close(zfi, ozfi) < RTOL
    # interior with the swapped vectors
    x = np.linalg.solve(ozh, ozfi)
    dyn2 = decks.Dynamic(dyn.title, dyn.l, dyn.dam, dyn.rho, dyn.freqs, kode, dyn.bc, dyn.cxi,
                         dyn.cyi, dyn.czi)
    ou, ot, ozd, ozs = orc.interec(mesh, dyn2, x, zdfi, 1.7, zmat, oconsts[0])
    u, t, zd, zs = ctx.interec(x, zdfi, kode, dyn.cxi, dyn.cyi, dyn.czi, zmat, 1.7, dyn.rho)
    np.testing.assert_array_equal(u, ou)
    np.testing.assert_array_equal(t, ot)
    assert _vec_close(zd, ozd) < RTOL and _vec_close(zs, ozs) < RTOL


def test_drop_in_static_diagonals_from_host(ctx, orc):
    """GHMATECD given HESTdiag/GESTdiag by the caller (the Fortran surface)."""
    mesh, dyn = decks.flat_patch_problem(6, 6, n_interior=0, freqs=(2.5,))
    ohd, ogd, oconsts = orc.ghmatece(mesh)
    ctx.send_shared_data(mesh)
    zmat = api.material(mesh.ge, mesh.rnu, dyn.dam, dyn.rho)
    zdfi = dyn.bc.astype(np.complex128)
    ctx.send_dyn_data(zdfi, dyn.kode)
    zh, zfi = ctx.ghmatecd(zmat, 2.5, consts=oconsts, hestdiag=ohd, gestdiag=ogd)
    ozh, ozfi, _ = orc.ghmatecd_accurate(mesh, ohd, ogd, oconsts, dyn.kode, zdfi, 2.5, zmat)
    assert close_blocks(zh, ozh, RTOL)[0]
    assert _vec_close(zfi, ozfi) < RTOL


def test_degenerate_element_reports_status(ctx):
    mesh, _ = decks.flat_patch_problem(4, 4, n_interior=0)
    c = mesh.cone.copy()
    c[5, 2] = c[5, 1]  # collapse element 6 into a triangle-with-a-zero-edge
    c[5, 3] = c[5, 0]
    bad = decks.Mesh(mesh.title, mesh.n, mesh.nbe, mesh.npg, mesh.ge, mesh.rnu, mesh.cx,
                     mesh.cy, mesh.cz, c)
    with pytest.raises(api.BemxError) as ei:
        ctx.send_shared_data(bad)
    assert ei.value.code == 1


def test_call_order_is_checked():
    c = api.Context(0)
    with pytest.raises(api.BemxError) as ei:
        c.ghmatece(1.0, 0.25) if c.mesh is not None else c._check(
            c._lib.bemx_ghmatece(c.h, 1.0, 0.25, None, None, None, None, None, None, None, None),
            "ghmatece")
    assert ei.value.code == 6
    c.close()
