"""The reference's general meshes (malhas.zip, the decks src/bench.sh:3 runs)
through the HIP path, against the oracle (GPU).

  * E510: the trench ("SOLO-TRINCHEIRA", 510 BE + 5 EE): a free surface with
    a trench cut into it, so most element pairs have dr/dn != 0, and 76
    interior points;
  * Esoil_1600: the soil box (1600 BE + 5 EE), 76 interior points.
Both at omega = 1, nu = 0.3, no damping, unit traction on the central
elements.  The decks are the reference's own files (tests/golden/).

Bars as for the configs: assembly 1e-12 per entry (block norm floor), the
FP64 solution 1e-12 and the mixed one 1e-10 relative against LAPACK on the
oracle's matrix (Oracle.ghmatecd_accurate), interior values 1e-12 from the
same boundary solution and 1e-10 along each path's own chain.  The driver
run on the trench is in test_gpu_driver.py.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, close_blocks
from bemx import api, decks

pytestmark = pytest.mark.gpu

RTOL = 1e-12
TOL_MIXED = 1e-10
DECKS = ("E510", "Esoil_1600")


def _rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / np.max(np.abs(b)))


def _deck(name):
    mesh = decks.read_e_deck(os.path.join(GOLDEN, "%se.dat" % name))
    dyn = decks.read_d_deck(os.path.join(GOLDEN, "%sd.dat" % name), mesh.nbe)
    return mesh, dyn


@pytest.fixture(scope="module")
def ctx():
    c = api.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("name", DECKS)
def test_general_mesh_assembly_solve_interior(ctx, orc, name):
    mesh, dyn = _deck(name)
    print("deck", name, "nbe", mesh.nbe, "L", dyn.l, flush=True)
    assert dyn.l == 76 and mesh.n == mesh.nbe + 5
    fr = float(dyn.freqs[0])
    zdfi = dyn.bc.astype(np.complex128)
    ctx.send_shared_data(mesh)
    hd, gd, consts = ctx.ghmatece(mesh.ge, mesh.rnu)
    ohd, ogd, oconsts = orc.ghmatece(mesh)
    assert close_blocks(hd.reshape(-1, 9), ohd.reshape(-1, 9), RTOL)[0]
    assert close_blocks(gd.reshape(-1, 9), ogd.reshape(-1, 9), RTOL)[0]
    zmat = api.material(mesh.ge, mesh.rnu, dyn.dam, dyn.rho)
    ctx.send_dyn_data(zdfi, dyn.kode)
    zh, zfi = ctx.ghmatecd(zmat, fr)
    ozh, ozfi, _ = orc.ghmatecd_accurate(mesh, ohd, ogd, oconsts, dyn.kode, zdfi, fr, zmat)
    ok, worst = close_blocks(zh, ozh, RTOL)
    print("ZH worst", worst, flush=True)
    assert ok, ("ZH", worst)
    assert _rel(zfi, ozfi) <= RTOL
    if name == "E510":
        # the trench has element pairs off the flat surface: dr/dn != 0, so
        # the normal-normal entries are nonzero off the diagonal
        rows = np.abs(ozh).reshape(mesh.nbe, 3, mesh.nbe, 3)
        assert np.count_nonzero(rows[:, 2, :, 2] > 1e-3 * rows.max()) > mesh.nbe
    x_ref = np.linalg.solve(ozh, ozfi)
    x64, _ = ctx.solve(zmat, fr, api.MODE_FP64)
    xm, st = ctx.solve(zmat, fr, api.MODE_MIXED)
    print("fp64 vs LAPACK %.2e, mixed vs LAPACK %.2e" % (_rel(x64, x_ref), _rel(xm, x_ref)),
          flush=True)
    assert _rel(x64, x_ref) <= RTOL
    assert _rel(xm, x_ref) <= TOL_MIXED and st[2] == 0.0
    # interior points: the same boundary solution on both sides, then each
    # path's own chain
    ou, ot, ozd, ozs = orc.interec(mesh, dyn, x_ref, zdfi, fr, zmat, oconsts[0])
    u, t, zd, zs = ctx.interec(x_ref, zdfi, dyn.kode, dyn.cxi, dyn.cyi, dyn.czi, zmat, fr,
                               dyn.rho)
    assert _rel(zd, ozd) <= RTOL and _rel(zs, ozs) <= RTOL
    u, t, zd, zs = ctx.interec(xm, zdfi, dyn.kode, dyn.cxi, dyn.cyi, dyn.czi, zmat, fr, dyn.rho)
    assert _rel(zd, ozd) <= TOL_MIXED and _rel(zs, ozs) <= TOL_MIXED
