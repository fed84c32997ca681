"""bemx_solve_sweep, the frequency loop of the reference's main program
(Main.for:167-200: GHMATECD + LINSOLVE per frequency) on the device (GPU).

Frequencies go in batches (8 by default): a batch is assembled two
frequencies per launch (k_dyn_offdiag_h2, the Gauss-point geometry shared)
on a second stream while the previous batch is solved, and the batch's
mixed solves share every launch of the factorisation and the refinement.
Each frequency's arithmetic is that of bemx_solve, so every solution must be
bit-identical to solving that frequency alone: five frequencies (one batch)
in both modes, and eleven (batches of six and five, the second assembled
beside the first one's solve), down to omega = 0.1 where the near-field series kernel carries
most of the element pairs near the diagonal, with two load cases.
"""
import numpy as np
import pytest

from bemx import api, decks

pytestmark = pytest.mark.gpu

FREQS = (0.1, 0.7, 2.0, 5.5, 9.0)
FREQS11 = (0.1, 0.2, 0.35, 0.5, 0.7, 1.0, 1.3, 2.0, 3.5, 5.5, 9.0)


@pytest.fixture(scope="module")
def ctx():
    mesh, dyn = decks.flat_patch_problem(24, 22, freqs=FREQS, n_interior=0, load_cases=2)
    c = api.Context(0)
    c.send_shared_data(mesh)
    c.ghmatece(mesh.ge, mesh.rnu)
    c.send_load_cases(dyn.load_cases.T.astype(np.complex128), dyn.kode)
    yield c, mesh, dyn
    c.close()


@pytest.mark.parametrize("mode,freqs", [(api.MODE_MIXED, FREQS), (api.MODE_FP64, FREQS),
                                        (api.MODE_MIXED, FREQS11)])
def test_sweep_equals_each_frequency_alone(ctx, mode, freqs):
    c, mesh, dyn = ctx
    zmat = api.material(mesh.ge, mesh.rnu, dyn.dam, dyn.rho)
    print("sweep of", len(freqs), "frequencies, mode", mode, flush=True)
    xs, st = c.solve_sweep(zmat, freqs, mode)
    assert xs.shape == (len(freqs), 3 * mesh.nbe, 2)
    assert np.all(st[:, 2] == 0.0)  # no fallback to FP64
    for k, fr in enumerate(freqs):
        x, s1 = c.solve(zmat, fr, mode)
        print("omega", fr, "max |sweep - alone|", float(np.max(np.abs(xs[k] - x))), flush=True)
        np.testing.assert_array_equal(xs[k], x)
        assert st[k, 0] == s1[0]  # same GMRES iterations


def test_sweep_captures_one_graph(ctx):
    """A fresh context's first mixed sweep of eleven frequencies (batches of
    six and five) captures one factorisation graph, the batch of six's; the
    other batch sizes its workspace holds are captured ahead with it, so the
    batch of five and a later single solve of the same mesh replay."""
    _, mesh, dyn = ctx
    c = api.Context(0)
    try:
        c.send_shared_data(mesh)
        c.ghmatece(mesh.ge, mesh.rnu)
        c.send_load_cases(dyn.load_cases.T.astype(np.complex128), dyn.kode)
        zmat = api.material(mesh.ge, mesh.rnu, dyn.dam, dyn.rho)
        c.solve_sweep(zmat, FREQS11, api.MODE_MIXED)
        s = c.last_stats()
        print("after the sweep", s, flush=True)
        assert (s["graph_captures"], s["graph_replays"]) == (1, 1), s
        c.solve(zmat, FREQS11[4], api.MODE_MIXED)
        s = c.last_stats()
        print("after one solve", s, flush=True)
        assert (s["graph_captures"], s["graph_replays"]) == (1, 2), s
    finally:
        c.close()


def test_sweep_refuses_per_case_kode(ctx):
    c, mesh, dyn = ctx
    zdfi = dyn.load_cases.T.astype(np.complex128)
    kode = np.repeat(np.asarray(dyn.kode, dtype=np.int32)[:, None], 2, axis=1)
    kode[0, 1] = 0  # the second case prescribes a displacement
    c.send_load_cases(zdfi, kode)
    try:
        zmat = api.material(mesh.ge, mesh.rnu, dyn.dam, dyn.rho)
        with pytest.raises(api.BemxError) as ei:
            c.solve_sweep(zmat, FREQS[:2], api.MODE_MIXED)
        assert ei.value.code == 3  # BEMX_ERR_ARG
    finally:
        c.send_load_cases(zdfi, dyn.kode)
