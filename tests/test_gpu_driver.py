"""The reference's own end-to-end tests (src/tests.sh:40-122), run through the
drop-in Fortran subroutine surface and its driver: the D output of each
reference deck must match the reference's golden result file under
dtester.c's rule (row 1-norms <= 1e-3, no NaN).  Also the KODE = 0 path of
the driver, the mixed-precision LINSOLVE through the same surface, and the
trench deck of malhas.zip.

The driver runs in a process of its own while this one holds the GPU: the
solver's cooperative launches need the GPU's cooperative-launch resource,
which the first process to use it keeps, so these tests come before any
test that solves in this process (the file order of the GPU suite)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, PKG
from bemx import decks, outputs

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(PKG, "bemx", "bemx_main")


def _lf_copy(src, dst):
    with open(src, "rb") as f:
        data = f.read().replace(b"\r\n", b"\n")
    with open(dst, "wb") as f:
        f.write(data)


def _run_driver(tmp_path, e, d, env_extra=None):
    eo, do = str(tmp_path / "out_e.txt"), str(tmp_path / "out_d.txt")
    env = dict(os.environ)
    env.update(env_extra or {})
    res = subprocess.run([DRIVER, e, d, eo, do], cwd=str(tmp_path), capture_output=True,
                         text=True, env=env, timeout=300)
    assert res.returncode == 0, res.stdout + res.stderr
    return open(do, "rb").read().decode("latin-1")


@pytest.mark.parametrize("size,solver", [(240, "fp64"), (2160, "fp64"), (240, "mixed")])
def test_reference_goldens_through_driver(tmp_path, size, solver):
    print("deck", size, solver, flush=True)
    e, d = str(tmp_path / "e.dat"), str(tmp_path / "d.dat")
    _lf_copy(os.path.join(GOLDEN, "ESOLO%dE_-5+5.DAT" % size), e)
    _lf_copy(os.path.join(GOLDEN, "ESOLO%dD_-5+5.DAT" % size), d)
    text = _run_driver(tmp_path, e, d, {"BEMX_SOLVER": solver})
    mesh = decks.read_e_deck(e)
    dyn = decks.read_d_deck(d, mesh.nbe)
    mine = outputs.parse_dynamic_results(text, mesh.nbe, dyn.l)
    gold = outputs.parse_dynamic_results(
        open(os.path.join(GOLDEN, "SSOLO%dD_-5+5.SOL" % size), "rb").read().decode("latin-1"),
        mesh.nbe, dyn.l)
    ok, err = outputs.compare_results(mine, gold)
    assert ok, err


def test_driver_matches_oracle_with_displacement_bcs(tmp_path, orc):
    """KODE = 0 on some DOFs and two frequencies: the driver's output equals
    the oracle chain to the printed 4 digits."""
    mesh, dyn = decks.flat_patch_problem(7, 6, n_interior=4, freqs=(0.7, 1.9))
    rng = np.random.default_rng(2)
    sel = rng.choice(3 * mesh.nbe, 10, replace=False)
    dyn.kode[sel] = 0
    dyn.bc[sel] = rng.normal(size=10) * 0.1
    e, d = str(tmp_path / "e.dat"), str(tmp_path / "d.dat")
    decks.write_e_deck(mesh, e)
    decks.write_d_deck(dyn, d)
    text = _run_driver(tmp_path, e, d)
    for blk in range(2):
        ref = orc.solve_problem(mesh, dyn, freq_index=blk, accurate=True)
        mine = outputs.parse_dynamic_results(text, mesh.nbe, dyn.l, block=blk)
        want = {"u": ref["u"].reshape(-1, 3), "t": ref["t"].reshape(-1, 3),
                "du": ref["zdsol"].reshape(-1, 3), "sigma": ref["zssol"].reshape(-1, 9)}
        # both sides printed/rounded to 4 significant digits
        back = outputs.parse_dynamic_results(
            outputs.format_frequency_block(ref["fr"], ref["u"], ref["t"], ref["zdsol"],
                                           ref["zssol"], mesh.nbe, dyn.l), mesh.nbe, dyn.l)
        for key in ("u", "t", "du", "sigma"):
            np.testing.assert_array_equal(mine[key], back[key])


def _echo_part(data: bytes) -> bytes:
    """A D output up to its first OUTPUTEC block (the INPUTECD echo)."""
    mark = b"\n\n " + b"*" * 79 + b"\n\n" + b" " * 30 + b"FREQ\xdc\xcaNCIA"
    k = data.find(mark)
    assert k > 0
    return data[:k]


@pytest.mark.parametrize("size", [240, 2160])
def test_output_files_in_the_reference_layout(tmp_path, size):
    """src/tests.sh:42-55: the E output is byte-identical to the reference's
    SSOLO*E file (INPUTECE's echo), the D output's INPUTECD echo is
    byte-identical to the SSOLO*D file's, and the D output carries the
    labels src/dtester.c:41-70,86-170 looks for."""
    e, d = str(tmp_path / "e.dat"), str(tmp_path / "d.dat")
    _lf_copy(os.path.join(GOLDEN, "ESOLO%dE_-5+5.DAT" % size), e)
    _lf_copy(os.path.join(GOLDEN, "ESOLO%dD_-5+5.DAT" % size), d)
    _run_driver(tmp_path, e, d)
    eo = open(str(tmp_path / "out_e.txt"), "rb").read()
    assert eo == open(os.path.join(GOLDEN, "SSOLO%dE_-5+5.SOL" % size), "rb").read()
    do = open(str(tmp_path / "out_d.txt"), "rb").read()
    gold = open(os.path.join(GOLDEN, "SSOLO%dD_-5+5.SOL" % size), "rb").read()
    assert _echo_part(do) == _echo_part(gold.replace(b"\r\n", b"\n"))
    for label in (b" N\xdaMERO DE ELEMENTOS DA MALHA=", b" N\xdaMERO DE ELEMENTOS DE CONTORNO=",
                  b" N\xdaMERO DE PONTOS INTERNOS=", b" N\xd3S DO CONTORNO", b"TRACTION X",
                  b"  PONTOS INTERNOS", b"SIGMAXX"):
        assert label in do, label


@pytest.mark.parametrize("solver", ["fp64", "mixed"])
def test_driver_on_the_trench(tmp_path, orc, solver):
    """The Fortran driver (GHMATECE, GHMATECD, LINSOLVE, INTEREC through the
    drop-in surface) on E510: its D output against the oracle chain's
    results printed in the same format, under dtester's rule."""
    e, d = str(tmp_path / "e.dat"), str(tmp_path / "d.dat")
    _lf_copy(os.path.join(GOLDEN, "E510e.dat"), e)
    _lf_copy(os.path.join(GOLDEN, "E510d.dat"), d)
    text = _run_driver(tmp_path, e, d, {"BEMX_SOLVER": solver})
    mesh = decks.read_e_deck(e)
    dyn = decks.read_d_deck(d, mesh.nbe)
    mine = outputs.parse_dynamic_results(text, mesh.nbe, dyn.l)
    ref = orc.solve_problem(mesh, dyn, accurate=True)
    want = outputs.parse_dynamic_results(
        outputs.format_frequency_block(ref["fr"], ref["u"], ref["t"], ref["zdsol"], ref["zssol"],
                                       mesh.nbe, dyn.l), mesh.nbe, dyn.l)
    ok, err = outputs.compare_results(mine, want)
    assert ok, err
    # tighter than dtester: every printed value (4 significant digits) within
    # one unit of its last digit, relative to the largest value of its row
    for key in ("u", "t", "du", "sigma"):
        a, b = mine[key], want[key]
        scale = np.max(np.abs(b), axis=1, keepdims=True)
        assert np.all(np.abs(a - b) <= 1.01e-3 * scale), key
