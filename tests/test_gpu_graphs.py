"""Factorisations run as captured HIP graphs (emfa_ctx_set_graphs): every
replay against the same launch sequence run eagerly (bitwise), with eager
work and synchronisations between the replays, and the refined answers of a
graph-factored batch against the oracle's FP64 solutions."""
import sys

import numpy as np
import pytest

from conftest import rel, sphere
from emfa_amd import hip
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _fills(ctx, M, freqs):
    Zs = []
    for f in freqs:
        Z = hip.Mat(ctx, M.n, M.n, hip.F32)
        hip.fill(ctx, M, Z, hip.CFIE, hip.wavenumber(f), 3, alpha=0.5)
        Zs.append(Z)
    return Zs


@pytest.fixture
def graph_ctx(ctx):
    ctx.lu_params(256, 0, 2)
    ctx.lu_update_mode(1)
    yield ctx
    ctx.graphs(False)
    ctx.lu_params(0, 0, -1)
    ctx.lu_update_mode(0)


@pytest.mark.parametrize("sub", [2, 3])
def test_graph_replays_equal_eager_batches(graph_ctx, sub):
    ctx = graph_ctx
    xyz, tri, rwg = sphere(sub, 1.0)
    M = hip.Mesh(ctx, xyz, tri, rwg)
    _say(f"graph batches n={M.n}")
    sets = [_fills(ctx, M, (2e8, 3e8, 4.5e8)), _fills(ctx, M, (2.5e8, 3.5e8, 5e8))]
    other = _fills(ctx, M, (4e8,))[0]
    try:
        ctx.graphs(False)
        eager = []
        for Zs in sets:
            lus = hip.lu_factor_batch(ctx, Zs, hip.F32)
            eager.append([l.factor().numpy() for l in lus])
            for l in lus:
                l.close()
        ctx.graphs(True)
        c0, l0 = ctx.graph_stats()
        # capture, replay, replay on other matrices, replay again; eager work between
        for rnd, i in enumerate((0, 0, 1, 0)):
            lus = hip.lu_factor_batch(ctx, sets[i], hip.F32)
            got = [l.factor().numpy() for l in lus]
            assert [l.pivoting() for l in lus] == [(0, 0)] * 3
            for z in range(3):
                assert np.array_equal(got[z], eager[i][z]), f"replay {rnd}, matrix {z}"
            for l in lus:
                l.close()
            ctx.graphs(False)
            lo = hip.LU(ctx, other, hip.F32)  # eager work between the replays
            ctx.sync()
            lo.close()
            ctx.graphs(True)
        c1, l1 = ctx.graph_stats()
        assert c1 - c0 == 1 and l1 - l0 == 4
    finally:
        for Z in sets[0] + sets[1] + [other]:
            Z.close()
        M.close()


def test_single_factor_through_a_graph(graph_ctx):
    ctx = graph_ctx
    rng = np.random.default_rng(5)
    n = 1100
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) + 4 * np.sqrt(n) * np.eye(n)
    Z = hip.Mat.from_numpy(ctx, A, hip.F32)
    ctx.graphs(False)
    le = hip.LU(ctx, Z, hip.F32)
    fe = le.factor().numpy()
    ctx.graphs(True)
    c0, _ = ctx.graph_stats()
    outs = []
    for _ in range(3):
        lg = hip.LU(ctx, Z, hip.F32)
        outs.append(lg.factor().numpy())
        assert lg.pivoting() == (0, 0)
        B = (rng.standard_normal((n, 7)) + 1j * rng.standard_normal((n, 7))).astype(np.complex64)
        X = hip.Mat.from_numpy(ctx, B, hip.F32)
        lg.solve(X)
        assert rel(X.numpy(), np.linalg.solve(A.astype(np.complex64).astype(np.complex128), B)) < 1e-4
        X.close()
        lg.close()
    assert ctx.graph_stats()[0] - c0 == 1
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])
    # a batch of one (graph) against the eager single factorisation: same kernels on other grids
    assert rel(outs[0], fe) < 1e-6
    le.close()
    Z.close()


def test_graph_factored_batch_refines_to_the_oracle(graph_ctx):
    """Three CFIE problems of one size factored together through a graph,
    refined in FP64 against the FP64 fill: the oracle's FP64 solutions."""
    ctx = graph_ctx
    ctx.graphs(True)
    xyz, tri, rwg = sphere(2, 1.0)
    M = hip.Mesh(ctx, xyz, tri, rwg)
    freqs = (2e8, 3e8, 4.5e8)
    Zs = _fills(ctx, M, freqs)
    lus = hip.lu_factor_batch(ctx, Zs, hip.F32)
    inc = hip.inc_table([(0.0, 0.0, 0.0), (60.0, 30.0, 90.0)])
    try:
        for f, lu in zip(freqs, lus):
            k = hip.wavenumber(f)
            Z64 = hip.Mat(ctx, M.n, M.n, hip.F64)
            hip.fill(ctx, M, Z64, hip.CFIE, k, 3, alpha=0.5)
            B = hip.Mat(ctx, M.n, 2, hip.F64)
            hip.rhs_fill(ctx, M, B, hip.CFIE, k, 0.5, inc)
            X = hip.Mat(ctx, M.n, 2, hip.F64)
            rep, ok = hip.refine(ctx, Z64, lu, B, X, tol=1e-12, maxit=10)
            assert ok
            Zo = O.fill_cfie(xyz, tri, rwg, k, 0.5, 3, 0)
            Xo = np.linalg.solve(Zo, O.rhs(xyz, tri, rwg, 1, k, 0.5, inc))
            _say(f"f={f:.2e}: refine iterations {rep.iterations}, rel to oracle {rel(X.numpy(), Xo):.2e}")
            assert rel(X.numpy(), Xo) < 1e-9
            for m in (Z64, B, X):
                m.close()
    finally:
        for m in lus + Zs:
            m.close()
        M.close()
