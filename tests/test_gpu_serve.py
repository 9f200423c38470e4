"""Solve server on the device: a request's answer is the same alone and in a
concurrent mix, FP32+refinement answers agree with the FP64 oracle within
0.1 dB, factors are reused, admission refuses what cannot fit and runs
within the budget what can."""
import numpy as np
import pytest

from conftest import rel
from emfa_amd import hip, serve, workload
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GIB = 1 << 30


def _req(kind, subdiv, axes, form, freq, nrhs, theta_max=180.0):
    alpha = 0.5 if form == hip.CFIE else 1.0
    th = np.linspace(0.0, theta_max, nrhs)
    return dict(kind=kind, subdiv=subdiv, axes=axes, formulation=form, alpha=alpha, freq=freq,
                angles=np.stack([th, 0 * th, 0 * th], 1))


def _submit(srv, r, prec):
    return srv.submit(r["angles"], kind=r["kind"], subdiv=r["subdiv"], axes=r["axes"], formulation=r["formulation"],
                      alpha=r["alpha"], freq=r["freq"], precision=prec)


TARGET = _req(serve.ICOSPHERE, 3, (0.6, 0.5, 0.4), hip.CFIE, 4.0e8, 9)
OTHERS = [
    _req(serve.ICOSPHERE, 2, (0.5, 0.5, 0.5), hip.EFIE, 3.0e8, 5),
    _req(serve.PLATE, 16, (1.0, 1.0, 1.0), hip.EFIE, 2.5e8, 12, 80.0),
    _req(serve.ICOSPHERE, 3, (0.9, 0.7, 0.5), hip.EFIE, 2.0e8, 33),
    _req(serve.PLATE, 24, (1.0, 1.0, 1.0), hip.EFIE, 5.0e8, 3, 80.0),
    _req(serve.ICOSPHERE, 2, (1.1, 0.4, 0.8), hip.CFIE, 5.5e8, 64),
    _req(serve.ICOSPHERE, 3, (0.6, 0.5, 0.4), hip.EFIE, 4.0e8, 2),
]


@pytest.mark.parametrize("prec", [hip.F32, hip.F64])
def test_result_same_alone_and_in_the_mix(prec):
    with serve.Server(workers=4, hbm_budget=32 * GIB, keep_current=True) as srv:
        alone = srv.wait(_submit(srv, TARGET, prec))
    assert alone["status"] == 0, alone["error"]
    assert not alone["cache_hit"]
    with serve.Server(workers=4, hbm_budget=32 * GIB, keep_current=True) as srv:
        ids = [_submit(srv, r, p) for r, p in zip(OTHERS[:3], [hip.F32, hip.F64, hip.F32])]
        tid = _submit(srv, TARGET, prec)
        ids += [_submit(srv, r, p) for r, p in zip(OTHERS[3:], [hip.F64, hip.F32, prec])]
        mixed = {}
        for _ in range(len(ids) + 1):
            res = srv.wait(-1)
            mixed[res["id"]] = res
        assert all(mixed[i]["status"] == 0 for i in ids + [tid]), [mixed[i]["error"] for i in ids + [tid]]
        again = srv.wait(_submit(srv, TARGET, prec))  # resident factor
    inmix = mixed[tid]
    assert again["cache_hit"]
    # deterministic kernels on one context's streams: bitwise the same
    for other in (inmix, again):
        assert np.array_equal(other["current"], alone["current"])
        assert np.array_equal(other["rcs_db"], alone["rcs_db"])
        assert other["refine_iterations"] == alone["refine_iterations"]


def test_fp32_refined_matches_fp64_oracle():
    r = _req(serve.ICOSPHERE, 2, (0.7, 0.5, 0.6), hip.EFIE, 3.5e8, 7)
    p = _req(serve.PLATE, 14, (1.0, 1.0, 1.0), hip.EFIE, 3.0e8, 5, 75.0)
    with serve.Server(workers=2, hbm_budget=8 * GIB, keep_current=True) as srv:
        got = [srv.wait(_submit(srv, x, hip.F32)) for x in (r, p)]
    for x, g in zip((r, p), got):
        assert g["status"] == 0, g["error"]
        assert g["relres"] <= 1e-8 and not g["fp64_fallback"]
        xyz, tri, rwg = workload._mesh(x)
        k = O.wavenumber(x["freq"])
        Z = O.fill_efie(xyz, tri, rwg, k, 3)
        inc = O.inc_table(x["angles"])
        V = O.rhs(xyz, tri, rwg, 0, k, 1.0, inc)
        X = np.linalg.solve(Z, V)
        ref = O.rcs(xyz, tri, rwg, k, X, -inc[:, 0:3], np.arange(inc.shape[0]))
        # the refinement solves the FP32-filled matrix to 1e-8 (below the
        # fill's own FP32 rounding, ~6e-8 relative); what remains is that rounding
        assert rel(g["current"], X) < 1e-5
        assert np.abs(g["rcs_db"] - ref).max() < 0.1
        assert np.abs(g["rcs_db"] - ref).max() < 1e-3


def test_factor_reuse_and_lru_eviction():
    a = _req(serve.ICOSPHERE, 3, (0.5, 0.5, 0.5), hip.EFIE, 3.0e8, 4)
    b = _req(serve.ICOSPHERE, 3, (0.8, 0.6, 0.5), hip.EFIE, 3.0e8, 4)
    n = serve.unknowns(serve.ICOSPHERE, 3)
    one = serve.footprint(n, 4, hip.F32)  # one request with its factor, nothing cached
    with serve.Server(workers=1, hbm_budget=one + one // 4) as srv:  # room for one factor only
        r1 = srv.wait(_submit(srv, a, hip.F32))
        r2 = srv.wait(_submit(srv, a, hip.F32))
        assert not r1["cache_hit"] and r2["cache_hit"]
        assert np.array_equal(r1["rcs_db"], r2["rcs_db"])
        assert srv.cached_factors() == 1
        r3 = srv.wait(_submit(srv, b, hip.F32))  # evicts a's factor
        assert r3["status"] == 0 and not r3["cache_hit"]
        assert srv.cached_factors() == 1
        r4 = srv.wait(_submit(srv, a, hip.F32))
        assert not r4["cache_hit"]
        assert np.array_equal(r1["rcs_db"], r4["rcs_db"])
        st = srv.stats()
        assert st["completed"] == 4 and st["cache_hits"] == 1
        assert st["peak_bytes"] <= one + one // 4


def test_admission_refuses_what_cannot_fit_and_serialises_what_can():
    big = _req(serve.ICOSPHERE, 3, (0.5, 0.5, 0.5), hip.EFIE, 3.0e8, 4)
    n = serve.unknowns(serve.ICOSPHERE, 3)
    need = serve.footprint(n, 4, hip.F64)
    with serve.Server(workers=3, hbm_budget=need // 2) as srv:
        res = srv.wait(_submit(srv, big, hip.F64))
        assert res["status"] == hip.EMFA_E_OOM
        assert "budget" in res["error"]
    # room for exactly one request at a time: three concurrent workers queue on memory
    reqs = [_req(serve.ICOSPHERE, 3, (0.5 + 0.1 * i, 0.5, 0.5), hip.EFIE, 3.0e8, 4) for i in range(3)]
    with serve.Server(workers=3, hbm_budget=need + need // 4) as srv:
        ids = [_submit(srv, r, hip.F64) for r in reqs]
        out = [srv.wait(i) for i in ids]
        assert all(o["status"] == 0 for o in out), [o["error"] for o in out]
        assert srv.stats()["peak_bytes"] <= need + need // 4


def test_mixed_stream_sample_served_matches_plain_fp64(ctx):
    reqs = [r for r in workload.mixed_stream(10, seed=11) if workload.unknowns(r) <= 8000][:6]
    plain, _, _ = workload.run_plain(ctx, reqs)
    with serve.Server(workers=3, hbm_budget=48 * GIB) as srv:
        served, st = workload.run_served(srv, reqs, hip.F32)
    assert st["completed"] == len(reqs) and st["failed"] == 0
    for p, s in zip(plain, served):
        assert np.abs(p - s["rcs_db"]).max() < 0.1


def test_hold_queues_a_batch_that_is_then_served_largest_first():
    import time
    small = _req(serve.ICOSPHERE, 2, (0.5, 0.5, 0.5), hip.EFIE, 3.0e8, 3)
    large = _req(serve.ICOSPHERE, 3, (0.7, 0.5, 0.5), hip.EFIE, 3.0e8, 3)
    with serve.Server(workers=1, hbm_budget=16 * GIB) as srv:
        srv.hold(True)
        ids = [_submit(srv, small, hip.F32), _submit(srv, large, hip.F32)]
        time.sleep(0.3)
        assert srv.stats()["completed"] == 0
        srv.hold(False)
        res = {i: srv.wait(i) for i in ids}
        assert all(r["status"] == 0 for r in res.values())
        # one worker: the larger new factor first although it was submitted second
        assert res[ids[1]]["queue_ms"] < res[ids[0]]["queue_ms"]
        # a hold is released by drain as well
        srv.hold(True)
        rid = _submit(srv, small, hip.F32)
        srv.drain()
        assert srv.wait(rid)["status"] == 0


def test_queued_requests_of_one_factor_are_batched_and_match_alone():
    # continuous batching: requests queued behind a busy worker that share a
    # factor are solved together (one stacked refinement); each answer stays
    # within the refinement tolerance of the same request served alone
    base = _req(serve.ICOSPHERE, 3, (0.7, 0.6, 0.5), hip.EFIE, 3.5e8, 5)
    mates = [dict(base, angles=np.stack([th, 0 * th, 0 * th], 1))
             for th in (np.linspace(0, 170, 5), np.linspace(10, 90, 9), np.array([33.0, 147.0]))]
    blocker = _req(serve.PLATE, 40, (1.0, 1.0, 1.0), hip.EFIE, 4.0e8, 3, 80.0)
    with serve.Server(workers=1, hbm_budget=16 * GIB, keep_current=True) as srv:
        alone = [srv.wait(_submit(srv, r, hip.F32)) for r in mates]
    with serve.Server(workers=1, hbm_budget=16 * GIB, keep_current=True) as srv:
        b = _submit(srv, blocker, hip.F32)
        ids = [_submit(srv, r, hip.F32) for r in mates]
        assert srv.wait(b)["status"] == 0
        got = [srv.wait(i) for i in ids]
        st = srv.stats()
    assert all(a["batch_size"] == 1 for a in alone)
    assert all(g["status"] == 0 and g["batch_size"] == 3 for g in got), [g["batch_size"] for g in got]
    assert st["batched"] == 2
    for a, g in zip(alone, got):
        assert g["nrhs"] == a["nrhs"] and g["relres"] <= 1e-8
        assert rel(g["current"], a["current"]) < 1e-6
        assert np.abs(g["rcs_db"] - a["rcs_db"]).max() < 1e-4


def test_new_factors_of_one_size_are_factored_together_and_match_alone(monkeypatch):
    # factor batching: requests for different new factors of one size queued
    # behind a busy worker are factored with one batched factorisation; each
    # answer equals the same request served with factor batching off
    reqs = [_req(serve.ICOSPHERE, 2, (0.8, 0.7, 0.6), form, f, 4)
            for form, f in ((hip.EFIE, 2.0e8), (hip.CFIE, 2.6e8), (hip.EFIE, 3.1e8), (hip.CFIE, 3.7e8))]
    blocker = _req(serve.PLATE, 40, (1.0, 1.0, 1.0), hip.EFIE, 4.0e8, 3, 80.0)
    with serve.Server(workers=1, hbm_budget=16 * GIB, keep_current=True) as srv:
        b = _submit(srv, blocker, hip.F32)
        ids = [_submit(srv, r, hip.F32) for r in reqs]
        assert srv.wait(b)["status"] == 0
        got = [srv.wait(i) for i in ids]
        st = srv.stats()
    assert st["factor_batches"] == 1 and st["batch_factors"] == 4, st
    assert all(g["status"] == 0 and not g["cache_hit"] and g["relres"] <= 1e-8 for g in got)
    monkeypatch.setenv("EMFA_SERVE_FACTOR_BATCH", "0")
    with serve.Server(workers=1, hbm_budget=16 * GIB, keep_current=True) as srv:
        alone = [srv.wait(_submit(srv, r, hip.F32)) for r in reqs]
        assert srv.stats()["factor_batches"] == 0
    for a, g in zip(alone, got):
        assert g["refine_iterations"] == a["refine_iterations"]
        assert rel(g["current"], a["current"]) < 1e-12
        assert np.abs(g["rcs_db"] - a["rcs_db"]).max() < 1e-9


def test_caller_mesh_served_like_the_same_synthetic_mesh():
    """A request carrying its own vertices and triangles (EMFA_MESH_USER) gets
    the same bits as the request naming the same mesh synthetically; a second
    request with equal mesh content reuses the resident factor (the key is a
    hash of the content)."""
    r = _req(serve.ICOSPHERE, 2, (0.7, 0.5, 0.6), hip.CFIE, 3.5e8, 6)
    xyz, tri = hip.icosphere(2, 0.7, 0.5, 0.6)
    with serve.Server(workers=2, hbm_budget=8 * GIB, keep_current=True) as srv:
        syn = srv.wait(_submit(srv, r, hip.F32))
        usr = srv.wait(srv.submit(r["angles"], formulation=r["formulation"], alpha=r["alpha"], freq=r["freq"],
                                  precision=hip.F32, mesh=(xyz, tri)))
        again = srv.wait(srv.submit(r["angles"][:2], formulation=r["formulation"], alpha=r["alpha"],
                                    freq=r["freq"], precision=hip.F32, mesh=(xyz.copy(), tri.copy())))
    for g in (syn, usr, again):
        assert g["status"] == 0, g["error"]
    assert not usr["cache_hit"] and again["cache_hit"]
    assert np.array_equal(usr["rcs_db"], syn["rcs_db"])
    assert np.array_equal(usr["current"], syn["current"])
    assert np.array_equal(again["rcs_db"], syn["rcs_db"][:2])


def test_mlfma_operator_resident_and_same_alone_and_in_the_mix():
    """solver=MLFMA keeps the multipole operator resident: the answer equals
    GMRES(30) on the same operator built directly (hip.MLFMA) bit for bit, a
    repeat reuses the operator, and the answer is the same in a mix of dense
    requests."""
    xyz, tri = hip.icosphere(3, 0.5)
    freq, alpha, box, tol = 1.2e9, 0.5, 0.25, 1e-5
    ang = np.array([[0.0, 0.0, 0.0], [60.0, 30.0, 0.0], [150.0, 0.0, 90.0]])
    ctx = hip.Context(0)
    try:
        rwg, _ = hip.rwg_build(tri)
        mesh = hip.Mesh(ctx, xyz, tri, rwg)
        k = hip.wavenumber(freq)
        op = hip.MLFMA(ctx, mesh, k, alpha, box)
        inc = hip.inc_table(ang)
        B = hip.Mat(ctx, mesh.n, 3, hip.F64)
        hip.rhs_fill(ctx, mesh, B, hip.CFIE, k, alpha, inc)
        X = hip.Mat(ctx, mesh.n, 3, hip.F64)
        rep, ok = op.gmres(B, X, m=30, tol=tol, max_restart=20)
        assert ok
        direct = hip.rcs(ctx, mesh, k, X, -inc[:, 0:3], np.arange(3, dtype=np.int32))
        xd = X.numpy()
        for m in (X, B, op, mesh):
            m.close()
    finally:
        ctx.close()
    kw = dict(formulation=hip.CFIE, alpha=alpha, freq=freq, mesh=(xyz, tri), solver=serve.MLFMA, mlfma_box=box,
              iter_tol=tol, iter_max=600)
    with serve.Server(workers=3, hbm_budget=16 * GIB, keep_current=True) as srv:
        alone = srv.wait(srv.submit(ang, **kw))
        ids = [_submit(srv, x, p) for x, p in zip(OTHERS[:2], [hip.F32, hip.F64])]
        tid = srv.submit(ang, **kw)
        ids.append(_submit(srv, OTHERS[2], hip.F32))
        got = {}
        for _ in range(len(ids) + 1):
            res = srv.wait(-1)
            got[res["id"]] = res
    assert alone["status"] == 0, alone["error"]
    assert not alone["cache_hit"] and got[tid]["cache_hit"]
    assert all(got[i]["status"] == 0 for i in ids + [tid])
    assert np.array_equal(alone["rcs_db"], direct)
    assert np.array_equal(alone["current"], xd)
    assert np.array_equal(got[tid]["rcs_db"], alone["rcs_db"])


def test_fastwt_operator_resident_and_same_as_the_direct_solve():
    """solver=FASTWT keeps the IP-SIE operator of a nonconformal caller mesh
    resident: the answer equals GMRES(30) on the same operator built
    directly (hip.FastWT) bit for bit, with RCS from its half-RWG far field,
    and a repeat reuses the operator."""
    xyz, tri = workload.nonconformal_cube((4, 5, 4, 5, 4, 5), 0.6)
    freq, alpha, box, tol = 3e8, 0.5, 0.2, 1e-6
    beta = 0.1 * abs(np.log10(0.07503))
    ang = np.array([[0.0, 0.0, 0.0], [70.0, 20.0, 90.0]])
    ctx = hip.Context(0)
    try:
        rwg, _ = hip.rwg_build(tri)
        mesh = hip.Mesh(ctx, xyz, tri, rwg)
        k = hip.wavenumber(freq)
        c = hip.contour_build(xyz, tri)
        op = hip.FastWT(ctx, mesh, c, k, alpha, box, beta)
        inc = hip.inc_table(ang)
        B = hip.Mat(ctx, c.n, 2, hip.F64)
        op.rhs(B, inc)
        X = hip.Mat(ctx, c.n, 2, hip.F64)
        rep, ok = op.gmres(B, X, m=30, tol=tol, max_restart=20)
        assert ok
        direct = np.array([op.rcs(X, -inc[j:j + 1, 0:3], col=j)[0] for j in range(2)])
        xd = X.numpy()
        for m in (X, B, op, mesh):
            m.close()
    finally:
        ctx.close()
    kw = dict(formulation=hip.CFIE, alpha=alpha, freq=freq, mesh=(xyz, tri), solver=serve.FASTWT, mlfma_box=box,
              iter_tol=tol, iter_max=600, fastwt_beta=beta)
    with serve.Server(workers=2, hbm_budget=16 * GIB, keep_current=True) as srv:
        first = srv.wait(srv.submit(ang, **kw))
        again = srv.wait(srv.submit(ang, **kw))
    assert first["status"] == 0, first["error"]
    assert not first["cache_hit"] and again["cache_hit"]
    assert first["unknowns"] == c.n
    assert np.array_equal(first["rcs_db"], direct)
    assert np.array_equal(first["current"], xd)
    assert np.array_equal(again["rcs_db"], first["rcs_db"])
