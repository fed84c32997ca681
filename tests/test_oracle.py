"""Pins of the CPU oracle (oracle/bem_oracle.c).

Each oracle function is checked three ways: against the reference's own
Fortran routines compiled under oracle/_ref, against the reference's golden
result decks, and against what the mathematics fixes (Legendre rules, the
Kelvin static limit, Hooke's law on finite-difference gradients, the wave
equation, a closed-form singular integral, LAPACK).
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, close_blocks
from bemx import decks, outputs
import oracle as oracle_mod

RNG = np.random.default_rng(7)
PI = 4.0 * np.arctan(1.0)


def material(orc, ge=1.0, rnu=0.25, dam=0.01, rho=1.0):
    return orc.material(ge, rnu, dam, rho)


def static_consts(ge, rnu):
    c1 = 1.0 / (16.0 * PI * ge * (1.0 - rnu))
    c2 = 3.0 - 4.0 * rnu
    c3 = -1.0 / (8.0 * PI * (1.0 - rnu))
    c4 = 1.0 - 2.0 * rnu
    return c1, c2, c3, c4


def rand_quad(center=(0, 0, 0), size=0.7):
    """A non-planar, non-parallelogram quad with positive orientation."""
    base = np.array([[-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]], float) * size / 2
    base += RNG.uniform(-0.08, 0.08, (4, 3)) * size
    return base + np.asarray(center, float)


# ------------------------------------------------------------------ GAULEG
@pytest.mark.parametrize("npg", [2, 3, 4, 6, 8, 10, 16])
def test_gauleg_matches_legendre(orc, npg):
    gi, ome = orc.gauleg(npg)
    x, w = np.polynomial.legendre.leggauss(npg)
    np.testing.assert_allclose(gi, x, rtol=0, atol=1e-14)  # GAULEG stops at |dz| <= 3e-14
    np.testing.assert_allclose(ome, w, rtol=0, atol=1e-14)


def test_gauleg_matches_reference(orc, ref):
    for npg in (4, 6, 8):
        a, b = orc.gauleg(npg), ref.gauleg(npg)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


# ------------------------------------------------------------------ point kernels vs reference
def test_point_kernels_match_reference(orc, ref):
    zmat = material(orc, 1.3, 0.27, 0.02, 1.1)
    # csqrt of glibc and of the Fortran runtime may differ in the last ulp
    np.testing.assert_allclose(zmat, ref.material(1.3, 0.27, 0.02, 1.1), rtol=1e-15)
    c1, c2, c3, c4 = static_consts(1.3, 0.27)
    for _ in range(20):
        xp = RNG.uniform(-2, 2, 3)
        xg = RNG.uniform(-2, 2, 3)
        n = RNG.normal(size=3)
        n /= np.linalg.norm(n)
        fr = RNG.uniform(0.1, 10.0)
        for f, args in (("solfund", (zmat[0], zmat[1], zmat[2], PI, fr)),
                        ("solfundif", (zmat[0], zmat[1], zmat[2], c1, c2, c3, c4, PI, fr))):
            a = getattr(orc, f)(xp, xg, n, *args)
            b = getattr(ref, f)(xp, xg, n, *args)
            for x, y in zip(a, b):
                ok, worst = close_blocks(x, y, 1e-12)
                assert ok, (f, worst)
        np.testing.assert_allclose(orc.solfune(xp, xg, n, c3, c4),
                                   ref.solfune(xp, xg, n, c3, c4), rtol=1e-14, atol=1e-16)


def test_element_quadratures_match_reference(orc, ref):
    zmat = material(orc)
    c1, c2, c3, c4 = static_consts(1.0, 0.25)
    for npg in (6, 8):
        gi, ome = orc.gauleg(npg)
        for _ in range(6):
            co = rand_quad(RNG.uniform(-1, 1, 3), 0.8)
            xp = RNG.uniform(-2, 2, 3) + np.array([0, 0, 1.5])
            eta = np.cross(co[1] - co[0], co[2] - co[0])
            eta /= np.linalg.norm(eta)
            np.testing.assert_allclose(orc.nonsinge(co, xp, eta, c3, c4, gi, ome),
                                       ref.nonsinge(co, xp, eta, c3, c4, gi, ome),
                                       rtol=1e-13, atol=1e-15)
            for x, y in zip(orc.nonsingd(co, xp, eta, *zmat, PI, 2.0, gi, ome),
                            ref.nonsingd(co, xp, eta, *zmat, PI, 2.0, gi, ome)):
                assert close_blocks(x, y, 1e-12)[0]
            for x, y in zip(orc.sigmaec(co, xp, eta, PI, 2.0, zmat[0], 1.0, zmat[1], zmat[2],
                                        gi, ome),
                            ref.sigmaec(co, xp, eta, PI, 2.0, zmat[0], 1.0, zmat[1], zmat[2],
                                        gi, ome)):
                assert close_blocks(x.reshape(3, 9), y.reshape(3, 9), 1e-12)[0]
            xm = co.mean(axis=0)
            np.testing.assert_allclose(orc.singge(co, xm, c1, c2, gi, ome),
                                       ref.singge(co, xm, c1, c2, gi, ome), rtol=1e-13,
                                       atol=1e-16)


# ------------------------------------------------------------------ whole routines vs reference
def _deck(size):
    mesh = decks.read_e_deck(os.path.join(GOLDEN, "ESOLO%dE_-5+5.DAT" % size))
    dyn = decks.read_d_deck(os.path.join(GOLDEN, "ESOLO%dD_-5+5.DAT" % size), mesh.nbe)
    return mesh, dyn


@pytest.mark.parametrize("case", ["esolo240", "c1"])
def test_chain_matches_reference(orc, ref, case):
    if case == "esolo240":
        mesh, dyn = _deck(240)
    else:
        mesh, dyn = decks.config_problem("c1")
    a = orc.solve_problem(mesh, dyn)
    b = ref.solve_problem(mesh, dyn)
    np.testing.assert_array_equal(orc.normvec(mesh), ref.normvec(mesh))
    np.testing.assert_allclose(a["consts"], b["consts"], rtol=1e-15)
    for k in ("hestdiag", "gestdiag"):
        ok, worst = close_blocks(a[k].reshape(-1, 9), b[k].reshape(-1, 9), 1e-12)
        assert ok, (k, worst)
    ok, worst = close_blocks(a["zh"], b["zh"], 1e-12)
    assert ok, ("zh", worst)
    for k in ("zfi", "x", "zdsol", "zssol"):
        rel = np.max(np.abs(a[k] - b[k])) / np.max(np.abs(b[k]))
        assert rel < 1e-12, (k, rel)


# ------------------------------------------------------------------ reference goldens
@pytest.mark.parametrize("size", [240, 2160])
def test_oracle_reproduces_reference_goldens(orc, size):
    mesh, dyn = _deck(size)
    o = orc.solve_problem(mesh, dyn)
    gold = outputs.parse_dynamic_results(
        open(os.path.join(GOLDEN, "SSOLO%dD_-5+5.SOL" % size), "rb").read().decode("latin-1"),
        mesh.nbe, dyn.l)
    mine = {"u": o["u"].reshape(-1, 3), "t": o["t"].reshape(-1, 3),
            "du": o["zdsol"].reshape(-1, 3), "sigma": o["zssol"].reshape(-1, 9)}
    ok, err = outputs.compare_results(mine, gold)
    assert ok, err


# ------------------------------------------------------------------ dtester's rule
DTESTER = os.path.join(ROOT, "oracle", "_ref", "dtester")
_DT_LABELS = {"Error in boundary nodes": "u", "Errors in tractions": "t",
              "Errors in Internal Displacements": "du", "Errors in Internal Sigmas": "sigma"}


def _run_dtester(path, path_sol):
    r = subprocess.run([DTESTER, path, path_sol], capture_output=True, timeout=60)
    err = {}
    for line in r.stdout.decode("latin-1").splitlines():
        label, _, val = line.partition(":")
        if label in _DT_LABELS:
            err[_DT_LABELS[label]] = float(val)
    return r.returncode, err


@pytest.mark.skipif(not os.path.exists(DTESTER),
                    reason="oracle/_ref/dtester not built (make -C oracle dtester)")
def test_compare_results_is_dtesters_rule(tmp_path):
    """outputs.compare_results against the reference's own checker
    (src/dtester.c, built from source by the oracle recipe) on copies of the
    SSOLO240D golden with differences planted in its results block: equal
    errors per section and the same verdict.  The plants separate a row
    maximum from a sum (two rows of u differ), and the stresses' per-point
    largest row from a sum over all nine entries."""
    mesh, dyn = _deck(240)
    nbe, l = mesh.nbe, dyn.l
    gold_path = os.path.join(GOLDEN, "SSOLO240D_-5+5.SOL")
    text = open(gold_path, "rb").read().decode("latin-1").replace("\r\n", "\n")
    gold = outputs.parse_dynamic_results(text, nbe, l)
    lines = text.split("\n")
    banner = next(i for i, s in enumerate(lines) if "NCIA =" in s)
    head = "\n".join(lines[:banner - 3]) + "\n"
    fr = float(lines[banner].split("=")[1].replace("D", "E"))

    def write(res, name):
        body = outputs.format_frequency_block(fr, res["u"].ravel(), res["t"].ravel(), res["du"].ravel(),
                                              res["sigma"].ravel(), nbe, l)
        p = str(tmp_path / name)
        open(p, "wb").write((head + body).encode("latin-1"))
        return p

    def plant(res, key, row, col, dz):
        a = res[key]
        a[row, col] += dz * max(1.0, abs(a[row, col]))

    cases = []
    same = {k: v.copy() for k, v in gold.items()}
    cases.append(("unchanged", same))
    small = {k: v.copy() for k, v in gold.items()}
    plant(small, "u", 4, 1, 3e-4)
    plant(small, "u", 8, 0, 2e-4j)
    plant(small, "du", 1, 2, 4e-4)
    plant(small, "sigma", 0, 0, 1e-4)
    plant(small, "sigma", 0, 1, 2e-4)
    plant(small, "sigma", 0, 8, 2.5e-4)
    plant(small, "sigma", 3, 4, 1e-4)
    cases.append(("within", small))
    big = {k: v.copy() for k, v in small.items()}
    plant(big, "t", 2, 0, 8e-4)
    plant(big, "t", 2, 2, -8e-4j)
    plant(big, "sigma", 5, 3, 7e-4)
    cases.append(("beyond", big))

    verdicts = []
    for name, res in cases:
        p = write(res, name + ".sol")
        rc, dt = _run_dtester(p, gold_path)
        mine = outputs.parse_dynamic_results(open(p, "rb").read().decode("latin-1"), nbe, l)
        ok, err = outputs.compare_results(mine, gold)
        assert set(dt) == set(err), (name, dt)
        for k in err:
            assert abs(err[k] - dt[k]) <= 1e-6 * max(dt[k], 1e-12), (name, k, err[k], dt[k])
        assert ok == (rc == 0), (name, rc, err)
        verdicts.append(ok)
    assert verdicts == [True, True, False]


# ------------------------------------------------------------------ mathematics
def test_quadrature_rule_area_and_centroid(orc):
    """Sum of OME*OME*DET is the element area, and the first moment is the
    area-weighted centroid (pins the shape functions, Jacobian and weights)."""
    gi, ome = orc.gauleg(8)
    co = np.array([[0, 0, 0], [2.0, 0, 0], [2.5, 1.5, 0], [0.3, 1.0, 0]])
    xg, w = oracle_mod.quad_rule(orc, co, gi, ome)
    # shoelace area and centroid of the planar quad
    x, y = co[:, 0], co[:, 1]
    cross = x * np.roll(y, -1) - np.roll(x, -1) * y
    area = 0.5 * cross.sum()
    cxa = ((x + np.roll(x, -1)) * cross).sum() / (6 * area)
    cya = ((y + np.roll(y, -1)) * cross).sum() / (6 * area)
    assert abs(w.sum() - area) < 1e-13
    assert abs((w * xg[:, 0]).sum() / area - cxa) < 1e-13
    assert abs((w * xg[:, 1]).sum() / area - cya) < 1e-13


def test_dynamic_kernel_static_limit(orc):
    """As omega -> 0 the dynamic U, T tend to the Kelvin static kernels, with
    a deviation of order k r."""
    ge, rnu = 1.0, 0.25
    zmat = orc.material(ge, rnu, 0.0, 1.0)
    c1, c2, c3, c4 = static_consts(ge, rnu)
    xp = np.array([0.1, -0.2, 0.05])
    xg = np.array([0.9, 0.4, -0.3])
    n = np.array([0.3, -0.2, 0.93])
    n /= np.linalg.norm(n)
    r = np.linalg.norm(xg - xp)
    rd = (xg - xp) / r
    ust = (c1 / r) * (c2 * np.eye(3) + np.outer(rd, rd))
    tst = orc.solfune(xp, xg, n, c3, c4)
    prev = None
    for fr in (1e-1, 1e-2):
        zu, zt = orc.solfund(xp, xg, n, *zmat, PI, fr)
        du = np.max(np.abs(zu - ust)) / np.max(np.abs(ust))
        dt = np.max(np.abs(zt - tst)) / np.max(np.abs(tst))
        assert du < 3 * fr * r and dt < 3 * fr * r, (fr, du, dt)
        if prev is not None:  # first-order convergence
            assert du < prev[0] / 5 and dt < prev[1] / 5
        prev = (du, dt)


def _hooke_traction(ufun, x, n, lam, mu, h=1e-5):
    """traction_j = sigma_jm n_m of the displacement field ufun(x)[j]."""
    g = np.zeros((3, 3), dtype=complex)  # g[j, m] = d u_j / d x_m
    for m in range(3):
        e = np.zeros(3)
        e[m] = h
        g[:, m] = (ufun(x + e) - ufun(x - e)) / (2 * h)
    sig = lam * np.trace(g) * np.eye(3) + mu * (g + g.T)
    return sig @ n


@pytest.mark.parametrize("fr", [0.0, 1.5])
def test_traction_kernel_is_hooke_of_displacement_kernel(orc, fr):
    """T_ij (traction j at the field point, load i) equals Hooke's law applied
    to the displacement field U_i.(x); pins psi', chi', A, B, C and T's
    index pattern independently of their formulas."""
    ge, rnu, rho = 1.0, 0.25, 1.0
    zmat = orc.material(ge, rnu, 0.02, rho)
    c1, c2, c3, c4 = static_consts(ge, rnu)
    xp = np.array([0.0, 0.0, 0.0])
    xg = np.array([0.7, -0.4, 0.5])
    n = np.array([0.2, 0.5, -0.84])
    n /= np.linalg.norm(n)
    if fr == 0.0:
        lam = 2 * ge * rnu / (1 - 2 * rnu)
        mu = ge

        def ucol(i):
            def f(x):
                r = np.linalg.norm(x - xp)
                rd = (x - xp) / r
                return (c1 / r) * (c2 * np.eye(3)[i] + rd[i] * rd)
            return f
        t = orc.solfune(xp, xg, n, c3, c4)
    else:
        mu = zmat[0]
        lam = zmat[2] ** 2 * rho - 2 * mu

        def ucol(i):
            return lambda x: orc.solfund(xp, x, n, *zmat, PI, fr)[0][i, :]
        t = orc.solfund(xp, xg, n, *zmat, PI, fr)[1]
    for i in range(3):
        tt = _hooke_traction(ucol(i), xg, n, lam, mu)
        np.testing.assert_allclose(t[i, :], tt, rtol=0, atol=1e-7 * np.abs(t).max())


def test_displacement_kernel_solves_navier_equation(orc):
    """mu lap u + (lam+mu) grad div u + rho w^2 u = 0 away from the source, and
    the waves decay (outgoing with damping)."""
    ge, rnu, rho, fr = 1.0, 0.25, 1.0, 2.0
    zmat = orc.material(ge, rnu, 0.01, rho)
    mu = zmat[0]
    lam = zmat[2] ** 2 * rho - 2 * mu
    xp = np.zeros(3)
    x0 = np.array([0.6, 0.3, -0.5])
    n = np.array([0.0, 0.0, 1.0])
    h = 2e-3

    def u(x, i):
        return orc.solfund(xp, x, n, *zmat, PI, fr)[0][i, :]

    for i in range(3):
        hess = np.zeros((3, 3, 3), dtype=complex)  # hess[j, a, b]
        for a in range(3):
            for b in range(3):
                ea = np.eye(3)[a] * h
                eb = np.eye(3)[b] * h
                hess[:, a, b] = (u(x0 + ea + eb, i) - u(x0 + ea - eb, i) - u(x0 - ea + eb, i)
                                 + u(x0 - ea - eb, i)) / (4 * h * h)
        lap = np.array([np.trace(hess[j]) for j in range(3)])
        graddiv = np.array([sum(hess[j, j, a] for j in range(3)) for a in range(3)])
        res = mu * lap + (lam + mu) * graddiv + rho * fr ** 2 * u(x0, i)
        scale = np.abs(mu * lap).max()
        assert np.abs(res).max() < 1e-4 * scale, (i, np.abs(res).max(), scale)
    # damping: |U| at 2r is less than |U| at r / 2 (1/r spreading alone)
    r1 = np.abs(orc.solfund(xp, x0, n, *zmat, PI, fr)[0]).max()
    r2 = np.abs(orc.solfund(xp, 2 * x0, n, *zmat, PI, fr)[0]).max()
    assert r2 < r1


def test_singular_static_g_closed_form(orc):
    """GESTdiag of a flat rectangle equals the closed-form integral of the
    Kelvin U over the rectangle (collocation at the centre)."""
    ge, rnu = 1.0, 0.25
    c1, c2, c3, c4 = static_consts(ge, rnu)
    p, q = 0.3125, 0.3333333333333333
    co = np.array([[-p, -q, 0], [p, -q, 0], [p, q, 0], [-p, q, 0]])
    i0 = 4 * (p * np.arcsinh(q / p) + q * np.arcsinh(p / q))
    ixx = 4 * q * np.arcsinh(p / q)
    iyy = 4 * p * np.arcsinh(q / p)
    expect = c1 * np.diag([c2 * i0 + ixx, c2 * i0 + iyy, c2 * i0])
    # the collapsed-quad rule is a quadrature: 8x8 points carry ~1e-6 of
    # error, and it converges to the closed form as NPG grows
    for npg, tol in ((8, 3e-6), (16, 1e-10)):
        gi, ome = orc.gauleg(npg)
        g = orc.singge(co, np.zeros(3), c1, c2, gi, ome)
        np.testing.assert_allclose(g, expect, rtol=0, atol=tol * np.abs(expect).max())


def test_stress_kernels_are_hooke_of_interior_displacement(orc):
    """ZD(k,i,j) and ZS(k,i,j) are the stresses of the interior displacement
    fields G(x)[:,k] and -H(x)[:,k] (finite-difference gradient in x)."""
    ge, rnu, rho, fr = 1.0, 0.25, 1.0, 1.3
    zmat = orc.material(ge, rnu, 0.01, rho)
    mu = zmat[0]
    lam = zmat[2] ** 2 * rho - 2 * mu
    gi, ome = orc.gauleg(8)
    co = rand_quad((0.2, -0.1, 0.0), 0.9)
    eta = np.cross(co[1] - co[0], co[2] - co[0])
    eta /= np.linalg.norm(eta)
    x0 = np.array([0.5, 0.4, -0.9])
    zd, zs = orc.sigmaec(co, x0, eta, PI, fr, zmat[0], rho, zmat[1], zmat[2], gi, ome)
    h = 1e-5
    for which, kern in (("G", zd), ("H", zs)):
        for k in range(3):
            grad = np.zeros((3, 3), dtype=complex)  # grad[a, m] = d u_a / d x_m
            for m in range(3):
                e = np.eye(3)[m] * h
                hp, gp = orc.nonsingd(co, x0 + e, eta, *zmat, PI, fr, gi, ome)
                hm, gm = orc.nonsingd(co, x0 - e, eta, *zmat, PI, fr, gi, ome)
                mp = gp if which == "G" else hp
                mm = gm if which == "G" else hm
                grad[:, m] = (mp[:, k] - mm[:, k]) / (2 * h)
            sig = lam * np.trace(grad) * np.eye(3) + mu * (grad + grad.T)
            np.testing.assert_allclose(kern[k], sig, rtol=0, atol=1e-6 * np.abs(kern).max())


def test_static_h_rows_and_free_term(orc):
    """RIGID_BODY makes each static H block row sum to zero over all N
    elements; on the smooth interior of the patch the free term is 1/2 I."""
    mesh, _ = decks.config_problem("c1")
    hd, gd, consts = orc.ghmatece(mesh)
    nx = 16
    centre = 7 + nx * 7
    np.testing.assert_allclose(hd[centre], 0.5 * np.eye(3), atol=2e-3)
    assert np.all(np.linalg.eigvalsh(gd[centre]) > 0)  # G diagonal block is SPD


def test_linsolve_lapack_matches_plain_lu(orc):
    mesh, dyn = decks.config_problem("c1")
    o = orc.solve_problem(mesh, dyn)
    x2 = orc.zgesv_plain(o["zh"], o["zfi"])
    assert np.max(np.abs(o["x"] - x2)) < 1e-13 * np.max(np.abs(o["x"]))
    res = np.linalg.norm(o["zh"] @ o["x"] - o["zfi"]) / (
        np.linalg.norm(o["zh"], 1) * np.linalg.norm(o["x"]))
    assert res < 1e-15


def test_kode_swap_round_trip(orc):
    """Prescribing on some DOFs the displacement that the all-traction
    problem produced (in the reference's scaled form u / mu*) returns the
    traction there, and INTEREC's un-swap restores u / mu* and t mu*."""
    mesh, dyn = decks.flat_patch_problem(6, 5, n_interior=4)
    o = orc.solve_problem(mesh, dyn)
    u, t = o["x"], dyn.bc.astype(complex)
    zge = o["zmat"][0]
    sel = np.array([0, 4, 17, 31, 44])
    kode = dyn.kode.copy()
    kode[sel] = 0
    zdfi = t.copy()
    zdfi[sel] = u[sel] / zge
    hd, gd, consts = o["hestdiag"], o["gestdiag"], o["consts"]
    zh, zfi = orc.ghmatecd(mesh, hd, gd, consts, kode, zdfi, o["fr"], o["zmat"])
    x = np.linalg.solve(zh, zfi)
    np.testing.assert_allclose(x[sel], t[sel], atol=1e-12)
    keep = np.setdiff1d(np.arange(x.size), sel)
    np.testing.assert_allclose(x[keep], u[keep], rtol=0, atol=1e-12)
    dyn2 = decks.Dynamic(dyn.title, dyn.l, dyn.dam, dyn.rho, dyn.freqs, kode, dyn.bc,
                         dyn.cxi, dyn.cyi, dyn.czi)
    zfi2, zdfi2, zdsol, _ = orc.interec(mesh, dyn2, x, zdfi, o["fr"], o["zmat"], consts[0])
    np.testing.assert_allclose(zfi2[sel], u[sel] / zge, atol=1e-12)
    np.testing.assert_allclose(zdfi2[sel], t[sel] * zge, atol=1e-12)


def test_sampled_blocks_equal_the_full_ghmatecd(orc):
    """orc_ghmatecd_blocks forms each block with the arithmetic of the full
    GHMATECD restatement: the sampled blocks, diagonal ones included, are
    bit-identical to the full matrices (before the KODE swap)."""
    mesh, dyn = decks.flat_patch_problem(9, 8, n_interior=0, freqs=(0.7,), dam=0.02)
    hd, gd, consts = orc.ghmatece(mesh)
    zmat = orc.material(mesh.ge, mesh.rnu, dyn.dam, dyn.rho)
    zh, zfi, zg = orc.ghmatecd(mesh, hd, gd, consts, dyn.kode, dyn.bc.astype(np.complex128),
                               0.7, zmat, return_zg=True)
    rng = np.random.default_rng(3)
    bi = np.concatenate([np.arange(mesh.nbe), rng.integers(0, mesh.nbe, 60)])
    bj = np.concatenate([np.arange(mesh.nbe), rng.integers(0, mesh.nbe, 60)])
    hb, gb = orc.ghmatecd_blocks(mesh, hd, gd, consts, 0.7, zmat, bi, bj)
    for k, (i, j) in enumerate(zip(bi, bj)):
        assert np.array_equal(hb[k], zh[3 * i:3 * i + 3, 3 * j:3 * j + 3])
        assert np.array_equal(gb[k], zg[3 * i:3 * i + 3, 3 * j:3 * j + 3])


def _solfund_mp(xp, xg, rn, zge, zcs, zcp, pi, fr):
    """Solfund.for:29-60 in 40-digit arithmetic (mpmath), for pinning the
    extended-precision oracle where the FP64 formula loses digits."""
    import mpmath as mp
    mp.mp.dps = 40
    x = [mp.mpf(float(v)) for v in xg]
    p = [mp.mpf(float(v)) for v in xp]
    n = [mp.mpf(float(v)) for v in rn]
    zge, zcs, zcp = (mp.mpc(complex(v).real, complex(v).imag) for v in (zge, zcs, zcp))
    r = [x[i] - p[i] for i in range(3)]
    R = mp.sqrt(sum(v * v for v in r))
    rd = [v / R for v in r]
    drn = sum(rd[i] * n[i] for i in range(3))
    zwi = mp.mpc(0, fr)
    c0 = 1 / (4 * mp.pi * zge)
    c1, c2 = (zcp / zcs) ** 2, (zcs / zcp) ** 2
    zp, zs = -zwi / zcp * R, -zwi / zcs * R
    ep, es = mp.exp(zp), mp.exp(zs)
    fhi = (1 + 1 / zs ** 2 - 1 / zs) * es / R - c2 * (1 / zp ** 2 - 1 / zp) * ep / R
    cap = (1 + 3 / zs ** 2 - 3 / zs) * es / R - c2 * (1 + 3 / zp ** 2 - 3 / zp) * ep / R
    fhid = (-2 + zs + 3 / zs - 3 / zs ** 2) * es / R ** 2 - c2 * (-1 + 3 / zp - 3 / zp ** 2) * ep / R ** 2
    capd = (zs - 4 + 9 / zs - 9 / zs ** 2) * es / R ** 2 - c2 * (zp - 4 + 9 / zp - 9 / zp ** 2) * ep / R ** 2
    aa = fhid - cap / R
    bb = 4 * cap / R - 2 * capd
    cc = (c1 - 2) * (aa + bb / 2 - 3 * cap / R) - 2 * cap / R
    zu = np.zeros((3, 3), complex)
    zt = np.zeros((3, 3), complex)
    for i in range(3):
        for j in range(3):
            d = 1 if i == j else 0
            zu[i, j] = complex(c0 * (fhi * d - cap * rd[j] * rd[i]))
            zt[i, j] = complex((aa * (drn * d + rd[j] * n[i]) + rd[i] * rd[j] * drn * bb +
                                rd[i] * n[j] * cc) / (4 * mp.pi))
    return zu, zt


@pytest.mark.parametrize("kr", [1e-3, 1e-2, 0.1, 0.5, 2.0, 10.0])
def test_long_double_oracle_against_mpmath(orc, kr):
    """Both builds of the oracle lose digits like eps / (kR)^2 in Solfund
    (its 1/z^2 terms cancel between the waves), the x87 build with an eps
    2048 times smaller: it stays within 5e-18 / (kR)^2 relative of a
    40-digit evaluation (7e-13 at kR = 2.7e-3, the smallest kR of the
    benchmark meshes: the innermost Gauss point of a diagonal element at
    omega = 0.1), the FP64 build within 2e-14 / (kR)^2."""
    ld = oracle_mod.OracleLD()
    zmat = orc.material(1.3, 0.3, 0.02, 1.1)
    ks = abs(1.0 / zmat[1])  # |k_s| per unit frequency
    xp = np.array([0.3, -0.1, 0.05])
    rn = np.array([0.36, 0.48, 0.8])
    for fr in (0.1, 2.0):
        R = kr / (ks * fr)
        xg = xp + R * np.array([0.6, -0.64, 0.48])
        u_mp, t_mp = _solfund_mp(xp, xg, rn, *zmat, np.pi, fr)
        u_ld, t_ld = ld.solfund(xp, xg, rn, *zmat, np.pi, fr)
        u_64, t_64 = orc.solfund(xp, xg, rn, *zmat, np.pi, fr)
        for a, b in ((u_ld, u_mp), (t_ld, t_mp)):
            err = np.max(np.abs(a - b)) / np.max(np.abs(b))
            assert err <= max(2e-15, 5e-18 / kr ** 2), (kr, fr, err)
        err64 = np.max(np.abs(t_64 - t_mp)) / np.max(np.abs(t_mp))
        assert err64 <= max(2e-15, 2e-14 / kr ** 2), (kr, err64)


def test_long_double_oracle_matches_fp64_at_moderate_kr(orc):
    """Where the FP64 formulas are accurate (kR >~ 1 over most of the mesh),
    the extended-precision build agrees with the FP64 one to 1e-12 per
    entry: it is the same restatement."""
    ld = oracle_mod.OracleLD()
    mesh, dyn = decks.flat_patch_problem(7, 6, n_interior=0, freqs=(6.0,), dam=0.02)
    hd, gd, consts = orc.ghmatece(mesh)
    lhd, lgd, lconsts = ld.ghmatece(mesh)
    assert close_blocks(lhd.reshape(-1, 9), hd.reshape(-1, 9), 1e-13)[0]
    assert close_blocks(lgd.reshape(-1, 9), gd.reshape(-1, 9), 1e-13)[0]
    np.testing.assert_allclose(lconsts, consts, rtol=1e-15)
    zmat = orc.material(mesh.ge, mesh.rnu, dyn.dam, dyn.rho)
    zdfi = dyn.bc.astype(np.complex128)
    zh, zfi = orc.ghmatecd(mesh, hd, gd, consts, dyn.kode, zdfi, 6.0, zmat)
    lzh, lzfi = ld.ghmatecd(mesh, hd, gd, consts, dyn.kode, zdfi, 6.0, zmat)
    ok, worst = close_blocks(zh, lzh, 1e-12)
    assert ok, worst
    rng = np.random.default_rng(2)
    bi = np.concatenate([np.arange(mesh.nbe), rng.integers(0, mesh.nbe, 40)])
    bj = np.concatenate([np.arange(mesh.nbe), rng.integers(0, mesh.nbe, 40)])
    hb, gb = ld.ghmatecd_blocks(mesh, hd, gd, consts, 6.0, zmat, bi, bj)
    for k, (i, j) in enumerate(zip(bi, bj)):
        assert np.array_equal(hb[k], lzh[3 * i:3 * i + 3, 3 * j:3 * j + 3])
