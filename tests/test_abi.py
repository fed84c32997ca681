"""The C ABI library loads without a GPU and exports every entry point that
include/bemx.h declares; host-side logic runs without device calls."""
import ctypes as C
import os
import re

import numpy as np

from conftest import ROOT
from bemx import api, decks


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "bemx.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|double|const char \*)\s*(bemx_\w+)\s*\(", text, re.M)))


def test_header_and_python_agree():
    assert declared_symbols() == sorted(api.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = api.load_library()
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_gauleg_host_entry_point():
    """bemx_gauleg is host code: Legendre nodes and weights without a GPU."""
    for npg in (2, 6, 8, 12):
        gi, ome = api.gauleg(npg)
        x, w = np.polynomial.legendre.leggauss(npg)
        np.testing.assert_allclose(gi, x, atol=1e-14)
        np.testing.assert_allclose(ome, w, atol=1e-14)


def test_create_without_device_fails_cleanly():
    """With no GPU visible the context constructor returns an error code
    instead of aborting the process."""
    import pytest
    lib = api.load_library()
    h = C.c_void_p()
    code = lib.bemx_create(99, C.byref(h))
    assert code != 0 and not h.value
    with pytest.raises(api.BemxError):
        api.Context(99)


def test_material_matches_reference_formula(orc):
    for args in ((1.0, 0.25, 0.01, 1.0), (1.7, 0.31, 0.03, 2.2)):
        np.testing.assert_allclose(api.material(*args), orc.material(*args), rtol=1e-15)


def test_library_material_matches_reference_formula(orc):
    """bemx_material (host code of the library) against the oracle's
    restatement of Inputecd.for:58-60."""
    for args in ((1.0, 0.25, 0.01, 1.0), (1.7, 0.31, 0.03, 2.2), (0.5, 0.2, 0.005, 1.3)):
        np.testing.assert_allclose(api.material_c(*args), orc.material(*args), rtol=1e-15)


def test_request_footprint_and_server_without_device():
    """The admission bound covers at least ZH (16 nn^2 bytes) and, on the
    mixed path, its FP32 copy (8 nn^2), grows with every size, and is 0 for
    invalid sizes; a request loop cannot be created without a device."""
    import pytest
    lib = api.lib()
    for nx, ny in ((16, 15), (48, 45)):
        mesh = decks.flat_patch(nx, ny)
        nn = 3 * mesh.nbe
        fp = api.request_footprint(mesh, 1, api.MODE_MIXED)
        assert fp >= 16 * nn * nn + 8 * nn * nn
        assert api.request_footprint(mesh, 4, api.MODE_MIXED) > fp
        assert api.request_footprint(mesh, 1, api.MODE_FP64) < fp
        base = (mesh.nbe, mesh.n, mesh.np_, mesh.npg, 1, api.MODE_MIXED)
        for k in range(5):  # monotone in n, nbe, np, npg, nrhs
            grown = list(base)
            grown[k] += 1 if k != 1 else 0
            if k == 1:
                grown[1] += 7
            assert lib.bemx_request_footprint(*grown) >= lib.bemx_request_footprint(*base)
        assert lib.bemx_request_footprint(0, 5, 9, 8, 1, api.MODE_MIXED) == 0.0
        assert lib.bemx_request_footprint(mesh.nbe, mesh.n, mesh.np_, mesh.npg, 1, 7) == 0.0
    assert (api.request_footprint(decks.flat_patch(48, 45)) >
            api.request_footprint(decks.flat_patch(16, 15)))
    with pytest.raises(api.BemxError):
        api.Server(99, 2)


def test_deck_round_trip(tmp_path):
    mesh, dyn = decks.flat_patch_problem(5, 4, n_interior=3, freqs=(0.5, 2.0))
    pe, pd = str(tmp_path / "e.dat"), str(tmp_path / "d.dat")
    decks.write_e_deck(mesh, pe)
    decks.write_d_deck(dyn, pd)
    m2 = decks.read_e_deck(pe)
    d2 = decks.read_d_deck(pd, m2.nbe)
    assert (m2.n, m2.nbe, m2.npg) == (mesh.n, mesh.nbe, mesh.npg)
    np.testing.assert_array_equal(m2.cone, mesh.cone)
    np.testing.assert_array_equal(m2.cx, mesh.cx)
    np.testing.assert_array_equal(d2.freqs, dyn.freqs)
    np.testing.assert_array_equal(d2.kode, dyn.kode)
    np.testing.assert_array_equal(d2.bc, dyn.bc)
    np.testing.assert_array_equal(d2.czi, dyn.czi)


def test_reference_decks_parse():
    from conftest import GOLDEN
    mesh = decks.read_e_deck(os.path.join(GOLDEN, "ESOLO240E_-5+5.DAT"))
    dyn = decks.read_d_deck(os.path.join(GOLDEN, "ESOLO240D_-5+5.DAT"), mesh.nbe)
    assert (mesh.n, mesh.nbe, mesh.np_, mesh.npg) == (240, 100, 242, 8)
    assert (dyn.l, dyn.nfr, dyn.dam, dyn.rho) == (10, 1, 0.0, 1.0)
    assert dyn.bc.sum() == 4.0 and (dyn.kode == 1).all()


def test_generated_mesh_is_closed_and_outward():
    """Normals of the patch point +z and the enclosure's outward; the
    rigid-body free term of a closed surface is 1/2 I at a smooth point."""
    mesh, _ = decks.flat_patch_problem(6, 6, n_interior=0)
    import oracle
    et = oracle.Oracle("c").normvec(mesh)
    assert np.allclose(et[:mesh.nbe], [0, 0, 1])
    assert et[-1][2] < -0.99  # bottom points down
