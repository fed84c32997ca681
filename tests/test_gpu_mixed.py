"""Mixed-precision solve: FP32 LU with split-BF16x3 MFMA trailing update as
the preconditioner of FP64 GMRES-IR (GPU).

Bar: the refined solution agrees with the FP64 oracle's (LAPACK zgesv) to
1e-10 relative -- the accuracy contract of the mixed path -- without falling
back to the plain FP64 path, in a handful of GMRES iterations (a broken
BF16x3 product or pivot bookkeeping shows up as a useless preconditioner:
many iterations or a fallback).
"""
import os

import numpy as np
import pytest

from bemx import api, decks

pytestmark = pytest.mark.gpu

TOL = 1e-10


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


# 1100 = two full groups of eight 64-column blocks and a partial one: the
# update of everything beyond the next group (the bulk, on the second
# stream) is non-empty only from n > 1024.
@pytest.mark.parametrize("n", [5, 64, 65, 130, 257, 600, 1100])
def test_mixed_random(ctx, n):
    print("n", n, flush=True)
    a, b = _rand_system(n, 100 + n)
    x = ctx.linsolve(a, b, api.MODE_MIXED)
    st = ctx.last_stats()
    xr = np.linalg.solve(a, b)
    rel = np.max(np.abs(x - xr)) / np.max(np.abs(xr))
    assert rel < TOL, (rel, st)
    assert not st["fell_back"], st
    assert st["gmres_iters"] <= 8, st


def test_mixed_many_workgroups_matches_fp64(ctx):
    """8320 unknowns: the block-column kernel of every step then runs more
    workgroups (2 x 130) than the device has CUs, so some start only after
    others have finished.  The mixed solve must still converge in a few GMRES
    iterations to the FP64 path's solution (the GPU FP64 LU is the reference
    here: a CPU solve of this size would dominate the test)."""
    n = 8320
    a, b = _rand_system(n, 4242)
    x = ctx.linsolve(a.copy(order="F"), b, api.MODE_MIXED)
    st = ctx.last_stats()
    xf = ctx.linsolve(a.copy(order="F"), b, api.MODE_FP64)
    rel = np.max(np.abs(x - xf)) / np.max(np.abs(xf))
    assert rel < TOL, (rel, st)
    assert not st["fell_back"], st
    assert st["gmres_iters"] <= 8, st


def test_mixed_needs_in_block_pivoting(ctx):
    """Rows permuted inside each 64-row diagonal block: diagonal-block
    pivoting must find them."""
    n = 256
    rng = np.random.default_rng(3)
    a = 0.01 * (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)))
    for blk in range(n // 64):
        p = rng.permutation(64) + 64 * blk
        a[p, np.arange(64) + 64 * blk] += 5.0
    a = np.asfortranarray(a)
    b = rng.normal(size=n) + 1j * rng.normal(size=n)
    x = ctx.linsolve(a, b, api.MODE_MIXED)
    st = ctx.last_stats()
    xr = np.linalg.solve(a, b)
    assert np.max(np.abs(x - xr)) / np.max(np.abs(xr)) < TOL
    assert not st["fell_back"], st


def test_mixed_falls_back_or_converges_when_ill_conditioned(ctx):
    """cond ~ 1e9: the FP32 factors are poor; the answer must still meet the
    FP64 bar (GMRES converges or the FP64 path takes over)."""
    n = 200
    rng = np.random.default_rng(8)
    u, _ = np.linalg.qr(rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)))
    v, _ = np.linalg.qr(rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)))
    s = np.logspace(0, -9, n)
    a = np.asfortranarray((u * s) @ v.conj().T)
    xt = rng.normal(size=n) + 1j * rng.normal(size=n)
    b = a @ xt
    x = ctx.linsolve(a, b, api.MODE_MIXED)
    xr = np.linalg.solve(a, b)
    # both are limited by cond * eps; agree to that level
    assert np.max(np.abs(x - xr)) / np.max(np.abs(xr)) < 1e-5


@pytest.mark.parametrize("case", ["c1", "patch24"])
def test_mixed_bem_solve_matches_fp64_oracle(ctx, orc, case):
    if case == "c1":
        mesh, dyn = decks.config_problem("c1")
    else:
        mesh, dyn = decks.flat_patch_problem(24, 22, n_interior=0, freqs=(7.5,), dam=0.03)
    ref = orc.solve_problem(mesh, dyn, accurate=True)
    ctx.send_shared_data(mesh)
    ctx.ghmatece(mesh.ge, mesh.rnu)
    zmat = api.material(mesh.ge, mesh.rnu, dyn.dam, dyn.rho)
    ctx.send_dyn_data(dyn.bc.astype(np.complex128), dyn.kode)
    x, stats = ctx.solve(zmat, float(dyn.freqs[0]), api.MODE_MIXED)
    rel = np.max(np.abs(x - ref["x"])) / np.max(np.abs(ref["x"]))
    assert rel < TOL, rel
    assert stats[2] == 0.0  # no fallback


@pytest.mark.skipif(os.environ.get("BEMX_GRAPHS") == "0", reason="graphs switched off")
def test_graph_replays_match_and_recapture_on_size_change():
    """The mixed factorisation is captured into a graph once per (size,
    buffers) and replayed: replays of the same system, with eager work and a
    solve of another size (a re-capture) in between, give bitwise the same
    answer as the first (captured) run, and every one is within 1e-10 of the
    FP64 solution."""
    c = api.Context(0)
    try:
        a, b = _rand_system(300, 7)
        ref = np.linalg.solve(a, b)
        first = c.linsolve(a, b, api.MODE_MIXED)
        a2, b2 = _rand_system(200, 8)
        for rep in range(4):
            x = c.linsolve(a, b, api.MODE_MIXED)
            assert np.array_equal(x, first), rep
            assert np.max(np.abs(x - ref)) / np.max(np.abs(ref)) < TOL
            if rep == 1:
                y = c.linsolve(a2, b2, api.MODE_MIXED)  # other size: re-capture
                yr = np.linalg.solve(a2, b2)
                assert np.max(np.abs(y - yr)) / np.max(np.abs(yr)) < TOL
            c.linsolve(a, b, api.MODE_FP64)  # eager work between replays
    finally:
        c.close()
