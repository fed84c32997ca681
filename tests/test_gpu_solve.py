"""Parity of the HIP dense complex solve (LINSOLVE) with LAPACK zgesv (GPU).

Bar: the solution agrees with the oracle's (numpy/LAPACK zgesv on the same
matrix) to 1e-12 relative for well-conditioned systems, and the backward
error ||Ax-b|| / (||A|| ||x||) is at the FP64 level.  Odd sizes exercise the
last partial panel; the pivoting cases force row interchanges across panel
workgroups.
"""
import numpy as np
import pytest

from bemx import api, decks

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = api.Context(0)
    yield c
    c.close()


def _rand_system(n, seed, diag=4.0):
    rng = np.random.default_rng(seed)
    a = (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))) / np.sqrt(n)
    a += diag * np.eye(n)
    b = rng.normal(size=n) + 1j * rng.normal(size=n)
    return np.asfortranarray(a), b


def _check(a, b, x, tol=1e-12):
    xr = np.linalg.solve(a, b)
    rel = np.max(np.abs(x - xr)) / np.max(np.abs(xr))
    bw = np.linalg.norm(a @ x - b) / (np.linalg.norm(a, 1) * np.linalg.norm(x))
    assert rel < tol and bw < 1e-15 * max(1, a.shape[0] / 100), (rel, bw)


@pytest.mark.parametrize("n", [1, 3, 7, 64, 65, 127, 200, 333])
def test_random_sizes(ctx, n):
    print("n", n, flush=True)
    a, b = _rand_system(n, n)
    x = ctx.linsolve(a, b)
    _check(a, b, x)


@pytest.mark.parametrize("n", [150, 700])
def test_pivoting_across_workgroups(ctx, n):
    """Small diagonal, large entries far below it: every column pivots, and
    the pivot rows live in other panel workgroups."""
    rng = np.random.default_rng(5)
    a = 1e-3 * (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)))
    perm = rng.permutation(n)
    a[perm, np.arange(n)] += 10.0 + 5j
    a = np.asfortranarray(a)
    b = rng.normal(size=n) + 1j * rng.normal(size=n)
    x = ctx.linsolve(a, b)
    _check(a, b, x)


@pytest.mark.parametrize("rows", [256, 384])
def test_tall_panel_equals_plain_panel(ctx, rows, monkeypatch):
    """The FP64 panel's tall form (workgroups of 256 or 384 rows: the first
    128 in LDS, the rest updated in place in the matrix), which orders above
    32768 need, forced at an order the plain form also handles: every column
    pivots to a row far below it, in another workgroup or among a
    workgroup's rows kept in the matrix.  Both forms pick the same pivots and
    apply the same operations to every entry, so the solutions are
    bit-identical."""
    n = 3000
    rng = np.random.default_rng(21)
    a = 1e-3 * (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)))
    a[rng.permutation(n), np.arange(n)] += 10.0 + 5j
    a = np.asfortranarray(a)
    b = rng.normal(size=n) + 1j * rng.normal(size=n)
    print("plain panel, n", n, flush=True)
    x0 = ctx.linsolve(a, b)
    monkeypatch.setenv("BEMX_PANEL_ROWS", str(rows))
    print("tall panel, rows per workgroup", rows, flush=True)
    x1 = ctx.linsolve(a, b)
    np.testing.assert_array_equal(x1, x0)
    _check(a, b, x1)


def test_tied_pivot_candidates(ctx):
    """Equal |re|+|im| candidates: the lowest row wins, as IZAMAX."""
    n = 96
    a = np.zeros((n, n), dtype=np.complex128, order="F")
    for j in range(n):
        a[:, j] = 1.0 + 0.01 * j
        a[j, j] += 3.0
    a += np.diag(np.linspace(1, 2, n))
    b = np.arange(n) + 1j
    x = ctx.linsolve(a, b)
    _check(a, b, x, tol=1e-11)


def test_singular_matrix_reports_info(ctx):
    n = 80
    a, b = _rand_system(n, 9)
    a[:, 37] = 0.0
    with pytest.raises(api.BemxError) as ei:
        ctx.linsolve(a, b)
    assert ei.value.code == 2


@pytest.mark.parametrize("case", ["c1", "esolo240"])
def test_bem_chain_fp64(ctx, orc, case):
    """GHMATECE -> GHMATECD -> LINSOLVE -> INTEREC on the device against the
    oracle chain, with the device-resident ZH handed from GHMATECD to
    LINSOLVE as in the reference."""
    import os
    from conftest import GOLDEN
    if case == "c1":
        mesh, dyn = decks.config_problem("c1")
    else:
        mesh = decks.read_e_deck(os.path.join(GOLDEN, "ESOLO240E_-5+5.DAT"))
        dyn = decks.read_d_deck(os.path.join(GOLDEN, "ESOLO240D_-5+5.DAT"), mesh.nbe)
    res = api.solve_problem(ctx, mesh, dyn)
    ref = orc.solve_problem(mesh, dyn, accurate=True)
    for key in ("x", "zdsol", "zssol"):
        rel = np.max(np.abs(res[key] - ref[key])) / np.max(np.abs(ref[key]))
        assert rel < 1e-12, (key, rel)


def test_device_resident_solve_matches(ctx, orc):
    mesh, dyn = decks.flat_patch_problem(20, 19, n_interior=0, freqs=(2.0,))
    ctx.send_shared_data(mesh)
    ctx.ghmatece(mesh.ge, mesh.rnu)
    zmat = api.material(mesh.ge, mesh.rnu, dyn.dam, dyn.rho)
    ctx.send_dyn_data(dyn.bc.astype(np.complex128), dyn.kode)
    x, stats = ctx.solve(zmat, 2.0, api.MODE_FP64)
    ref = orc.solve_problem(mesh, dyn, accurate=True)
    rel = np.max(np.abs(x - ref["x"])) / np.max(np.abs(ref["x"]))
    assert rel < 1e-12, rel
    assert stats[3] > 0.0


@pytest.mark.parametrize("mode", [api.MODE_FP64, api.MODE_MIXED])
def test_load_cases_share_one_factorisation(ctx, mode):
    """Three load cases (random central traction patches, some KODE=0
    displacement columns) solved together match the same cases solved one at
    a time, column by column: one factorisation, several right-hand sides."""
    mesh, dyn = decks.flat_patch_problem(12, 11, n_interior=0, freqs=(1.5,), load_cases=3,
                                         seed=7)
    kode = dyn.kode.copy()
    kode[3 * 5:3 * 5 + 3] = 0      # element 6 fully prescribed displacement
    kode[3 * 40 + 1] = 0           # one mixed element
    cases = dyn.load_cases.T.astype(np.complex128)
    cases[kode == 0, :] = 0.01 * (1 + 1j)
    ctx.send_shared_data(mesh)
    ctx.ghmatece(mesh.ge, mesh.rnu)
    zmat = api.material(mesh.ge, mesh.rnu, dyn.dam, dyn.rho)
    ctx.send_load_cases(cases, kode)
    _, zfi_all = ctx.ghmatecd(zmat, 1.5, want_zh=False)
    x_all, _ = ctx.solve(zmat, 1.5, mode)
    assert x_all.shape == (3 * mesh.nbe, 3)
    for j in range(3):
        ctx.send_dyn_data(cases[:, j], kode)
        _, zfi = ctx.ghmatecd(zmat, 1.5, want_zh=False)
        assert np.max(np.abs(zfi_all[:, j] - zfi)) <= 1e-13 * np.max(np.abs(zfi))
        x, _ = ctx.solve(zmat, 1.5, mode)
        rel = np.max(np.abs(x_all[:, j] - x)) / np.max(np.abs(x))
        assert rel < (1e-13 if mode == api.MODE_FP64 else 1e-10), (j, rel)


@pytest.mark.parametrize("mode", [api.MODE_FP64, api.MODE_MIXED])
def test_load_cases_with_their_own_kode(ctx, orc, mode):
    """Three load cases, two KODE sets (cases 0 and 2 all traction, case 1
    with prescribed displacements on 15 DOFs): one column swap and
    factorisation per set, every case as if solved alone, and case 1 against
    the oracle's chain with its KODE (Ghmatecd.for:316-333)."""
    mesh, dyn = decks.flat_patch_problem(8, 7, n_interior=0, freqs=(1.7,), load_cases=3, seed=4)
    nn = 3 * mesh.nbe
    rng = np.random.default_rng(8)
    kode = np.ones((nn, 3), dtype=np.int32)
    sel = rng.choice(nn, 15, replace=False)
    kode[sel, 1] = 0
    zdfi = dyn.load_cases.T.astype(np.complex128).copy()
    zdfi[sel, 1] = rng.normal(size=15) + 1j * rng.normal(size=15)
    zmat = api.material(mesh.ge, mesh.rnu, dyn.dam, dyn.rho)
    ctx.send_shared_data(mesh)
    ctx.ghmatece(mesh.ge, mesh.rnu)
    ctx.send_load_cases(zdfi, kode)
    with pytest.raises(api.BemxError) as ei:
        ctx.ghmatecd(zmat, 1.7)
    assert ei.value.code == 3
    x, st = ctx.solve(zmat, 1.7, mode)
    tol = 1e-12 if mode == api.MODE_FP64 else 1e-10
    for j in range(3):
        ctx.send_load_cases(zdfi[:, j:j + 1], kode[:, j])
        xj, _ = ctx.solve(zmat, 1.7, mode)
        rel = np.max(np.abs(x[:, j] - xj)) / np.max(np.abs(xj))
        assert rel <= tol, (j, rel)
    hd, gd, consts = orc.ghmatece(mesh)
    ozh, ozfi, _ = orc.ghmatecd_accurate(mesh, hd, gd, consts, kode[:, 1], zdfi[:, 1], 1.7, zmat)
    xr = np.linalg.solve(ozh, ozfi)
    assert np.max(np.abs(x[:, 1] - xr)) / np.max(np.abs(xr)) <= tol
    # the same request through the request loop
    srv = api.Server(0, nslots=2, mode=mode)
    try:
        xs, info = srv.wait(srv.submit(mesh, kode, zdfi, (mesh.ge, mesh.rnu, dyn.dam, dyn.rho),
                                       1.7))
    finally:
        srv.close()
    assert np.max(np.abs(xs - x)) / np.max(np.abs(x)) <= tol
