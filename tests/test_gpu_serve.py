"""Request loop (bemx_server_*): continuous batching on one GPU.

Each request's result must be the same alone and in the mix: a request
solved through a context on its own, and the same request solved while
others of other sizes, frequencies and load-case counts run beside it on
other slots, agree to 1e-10 relative (the mixed path's bar; the arithmetic
of one request does not depend on its neighbours, so in practice they are
identical).  Admission: with a budget that holds one request at a time,
service intervals never overlap; a request larger than the whole budget is
refused.
"""
import numpy as np
import pytest

from bemx import api, decks

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _requests():
    """Six requests over two mesh sizes, with their own frequency, damping
    and 1..3 load cases (random central traction patches)."""
    rng = np.random.default_rng(5)
    out = []
    for i, (nx, ny) in enumerate([(8, 7), (12, 11), (8, 7), (12, 11), (10, 9), (8, 7)]):
        nrhs = 1 + i % 3
        mesh, dyn = decks.flat_patch_problem(nx, ny, n_interior=0, load_cases=nrhs, seed=i)
        fr = float(rng.uniform(0.1, 10.0))
        dam = float(rng.uniform(0.005, 0.05))
        zdfi = dyn.load_cases.T.astype(np.complex128)
        out.append((mesh, dyn.kode, zdfi, (mesh.ge, mesh.rnu, dam, dyn.rho), fr))
    return out


def _alone(req, mode):
    mesh, kode, zdfi, (ge, rnu, dam, rho), fr = req
    ctx = api.Context(0)
    try:
        ctx.send_shared_data(mesh)
        ctx.ghmatece(ge, rnu)
        ctx.send_load_cases(zdfi, kode)
        x, _ = ctx.solve(api.material(ge, rnu, dam, rho), fr, mode)
    finally:
        ctx.close()
    return x.reshape(zdfi.shape, order="F") if zdfi.shape[1] > 1 else x.reshape(-1, 1)


@pytest.mark.parametrize("mode", [api.MODE_MIXED, api.MODE_FP64])
def test_each_result_is_the_same_alone_and_in_the_mix(mode):
    reqs = _requests()
    alone = [_alone(r, mode) for r in reqs]
    srv = api.Server(0, nslots=3, mode=mode)
    try:
        ids = [srv.submit(m, k, z, p, f) for (m, k, z, p, f) in reqs]
        outs = [srv.wait(i) for i in ids]
    finally:
        srv.close()
    slots = set()
    for (x, info), ref in zip(outs, alone):
        x = x.reshape(ref.shape, order="F") if x.ndim == 1 else x
        rel = np.max(np.abs(x - ref)) / np.max(np.abs(ref))
        assert rel <= TOL, rel
        assert info["latency_ms"] >= info["service_ms"] > 0.0
        slots.add(int(info["slot"]))
    assert len(slots) > 1  # the work was spread over several slots


def test_mixed_request_matches_fp64_oracle(orc):
    """One request through the loop against the FP64 oracle (GHMATECE,
    GHMATECD, LAPACK zgesv)."""
    mesh, dyn = decks.flat_patch_problem(8, 7, n_interior=0, freqs=(3.0,))
    srv = api.Server(0, nslots=2, mode=api.MODE_MIXED)
    try:
        rid = srv.submit(mesh, dyn.kode, dyn.bc.astype(np.complex128),
                         (mesh.ge, mesh.rnu, dyn.dam, dyn.rho), 3.0)
        x, info = srv.wait(rid)
    finally:
        srv.close()
    ref = orc.solve_problem(mesh, dyn, accurate=True)["x"]
    assert np.max(np.abs(x - ref)) / np.max(np.abs(ref)) <= TOL
    assert info["fell_back"] == 0.0


def test_admission_holds_requests_to_the_budget():
    allr = _requests()
    # three requests on the 12 x 11 mesh (1 or 2 load cases, other frequencies)
    reqs = [allr[1], allr[3], allr[1][:4] + (0.7,)]
    fps = [api.request_footprint(r[0], r[2].shape[1], api.MODE_MIXED) for r in reqs]
    assert 2 * min(fps) > 1.5 * max(fps)
    srv = api.Server(0, nslots=3, hbm_budget=1.5 * max(fps), mode=api.MODE_MIXED)
    try:
        ids = [srv.submit(m, k, z, p, f) for (m, k, z, p, f) in reqs]
        infos = [srv.wait(i)[1] for i in ids]
        # any two of these requests together exceed the budget, so their
        # service intervals must not overlap
        spans = sorted((i["start_ms"], i["done_ms"]) for i in infos)
        for (s0, e0), (s1, e1) in zip(spans, spans[1:]):
            assert s1 >= e0, (spans,)
        mesh, kode, zdfi, p, f = reqs[0]
        small = api.Server(0, nslots=1, hbm_budget=0.5 * api.request_footprint(mesh, 1),
                           mode=api.MODE_MIXED)
        try:
            with pytest.raises(api.BemxError):
                small.submit(mesh, kode, zdfi[:, :1], p, f)
        finally:
            small.close()
    finally:
        srv.close()
