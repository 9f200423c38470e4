"""MLFMA operator on the device against the numpy restatement of the
reference algorithm (oracle/mlfma.py) and against the dense CFIE matrix.

The restatement pins the GPU operator tightly (same tree, samples,
interpolation, translation and patterns; FP32 on the device).  The dense
comparison pins the physics: for boxes two or more edges apart the
expansion reproduces the dense entries to ~1e-3 (checked in
tests/test_cpu_mlfma.py); the operator as a whole then sits within a few
per cent of the dense matrix on these coarse meshes, which is the
reference algorithm's own accuracy."""
import numpy as np
import pytest

from conftest import rel, sphere
from emfa_amd import hip
from oracle import mlfma as OM

pytestmark = pytest.mark.gpu

ALPHA = 0.5
CASES = {"s3_1.2GHz": (3, 0.5, 1.2e9, 0.25), "s4_800MHz": (4, 0.5, 8.0e8, 0.25)}


def _assert_rcs_close(r, ref):
    """Within 1 dB where the pattern is within 15 dB of its peak, 5 % in
    linear power overall (deep nulls amplify the operator's per-cent error)."""
    lit = ref > ref.max() - 15
    assert np.abs(r - ref)[lit].max() < 1.0
    p, q = 10 ** (r / 10), 10 ** (ref / 10)
    assert np.linalg.norm(p - q) / np.linalg.norm(q) < 5e-2


def _vec(n, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(n) + 1j * rng.standard_normal(n)


@pytest.fixture(scope="module")
def built(ctx):
    out = {}
    for name, (s, r, f, thr) in CASES.items():
        xyz, tri, rwg = sphere(s, r)
        k = hip.wavenumber(f)
        mesh = hip.Mesh(ctx, xyz, tri, rwg)
        op = hip.MLFMA(ctx, mesh, k, ALPHA, thr)
        tree = OM.Tree(xyz, tri, rwg, f, thr)
        out[name] = dict(xyz=xyz, tri=tri, rwg=rwg, k=k, mesh=mesh, op=op, tree=tree)
    yield out
    for v in out.values():
        v["op"].close()
        v["mesh"].close()


def _apply(ctx, op, x):
    n = x.shape[0]
    X = hip.Mat.from_numpy(ctx, x.reshape(n, 1), hip.F64)
    Y = hip.Mat(ctx, n, 1, hip.F64)
    op.matvec(X, Y)
    y = Y.numpy()[:, 0]
    X.close()
    Y.close()
    return y


def _dense(ctx, v):
    n = v["rwg"].shape[0]
    Z = hip.Mat(ctx, n, n, hip.F64)
    hip.fill(ctx, v["mesh"], Z, hip.CFIE, v["k"], 3, alpha=ALPHA)
    z = Z.numpy()
    Z.close()
    return z


@pytest.mark.parametrize("name", list(CASES))
def test_tree_and_samples_match_oracle(built, name):
    v = built[name]
    info, T = v["op"].info(), v["tree"]
    assert info["levels"] == T.level
    assert info["boxes"][1:] == [len(b) for b in T.boxes[1:]]
    assert info["multipoles"][2:] == [T.L[l] for l in range(2, T.level + 1)]
    rows, _ = T.near_pairs()
    assert info["near_nnz"] == len(rows)


@pytest.mark.parametrize("name", list(CASES))
def test_matvec_matches_oracle_restatement(ctx, built, name):
    v = built[name]
    T = v["tree"]
    n = v["rwg"].shape[0]
    x = _vec(n, 3)
    y = _apply(ctx, v["op"], x)
    Z = _dense(ctx, v)
    rows, cols = T.near_pairs()
    mask = np.zeros((n, n), dtype=bool)
    mask[rows, cols] = True
    ynear = (Z * mask) @ x
    rad, rec = T.rad_rec(v["xyz"], v["tri"], v["rwg"], ALPHA)
    yfar = T.far_matvec(x, rad, rec)
    assert rel(y, ynear + yfar) < 2e-5
    # the far part alone (FP32 aggregation / translation)
    assert rel(y - ynear, yfar) < 2e-4


def test_efie_operator_matches_oracle_restatement(ctx):
    # EMFA_EFIE: receive pattern = conjugated radiation pattern, EFIE near entries
    xyz, tri, rwg = sphere(3, 0.5)
    f = 1.2e9
    k = hip.wavenumber(f)
    mesh = hip.Mesh(ctx, xyz, tri, rwg)
    op = hip.MLFMA(ctx, mesh, k, 1.0, 0.25, kind=hip.EFIE)
    T = OM.Tree(xyz, tri, rwg, f, 0.25)
    n = rwg.shape[0]
    x = _vec(n, 11)
    y = _apply(ctx, op, x)
    Z = hip.Mat(ctx, n, n, hip.F64)
    hip.fill(ctx, mesh, Z, hip.EFIE, k, 3)
    rows, cols = T.near_pairs()
    mask = np.zeros((n, n), dtype=bool)
    mask[rows, cols] = True
    rad, rec = T.rad_rec(xyz, tri, rwg, 1.0)
    assert rel(y, (Z.numpy() * mask) @ x + T.far_matvec(x, rad, rec)) < 2e-5
    op.close()
    Z.close()
    mesh.close()


def test_matvec_close_to_dense(ctx, built):
    v = built["s4_800MHz"]
    n = v["rwg"].shape[0]
    x = _vec(n, 5)
    y = _apply(ctx, v["op"], x)
    assert rel(y, _dense(ctx, v) @ x) < 3e-2


def test_matvec_deterministic(ctx, built):
    v = built["s4_800MHz"]
    x = _vec(v["rwg"].shape[0], 7)
    assert np.array_equal(_apply(ctx, v["op"], x), _apply(ctx, v["op"], x))


def test_gmres_solution_and_rcs_near_dense_solve(ctx, built):
    v = built["s4_800MHz"]
    n, k = v["rwg"].shape[0], v["k"]
    inc = hip.inc_table([(0.0, 0.0, 0.0)])
    B = hip.Mat(ctx, n, 1, hip.F64)
    hip.rhs_fill(ctx, v["mesh"], B, hip.CFIE, k, ALPHA, inc)
    X = hip.Mat(ctx, n, 1, hip.F64)
    rep, ok = v["op"].gmres(B, X, m=30, tol=1e-4, max_restart=20)
    assert ok and rep.relres < 1e-4
    x_ml = X.numpy()[:, 0]
    Z = _dense(ctx, v)
    x_d = np.linalg.solve(Z, B.numpy()[:, 0])
    assert rel(x_ml, x_d) < 1e-1
    th = np.radians(np.arange(0, 181, 5.0))
    dirs = np.stack([np.sin(th), 0 * th, np.cos(th)], 1)
    Xd = hip.Mat.from_numpy(ctx, x_d.reshape(n, 1), hip.F64)
    r_ml = hip.rcs(ctx, v["mesh"], k, X, dirs, np.zeros(len(th), dtype=np.int32))
    r_d = hip.rcs(ctx, v["mesh"], k, Xd, dirs, np.zeros(len(th), dtype=np.int32))
    _assert_rcs_close(r_ml, r_d)
    for m in (B, X, Xd):
        m.close()


def test_bicgstab_on_the_operator_matches_gmres(ctx, built):
    """MLFMA::iterativeSolve's BiCGSTAB (MLFMA.cpp:402-452) on the device
    operator reaches the operator's own GMRES solution."""
    v = built["s4_800MHz"]
    n, k = v["rwg"].shape[0], v["k"]
    inc = hip.inc_table([(0.0, 0.0, 0.0), (90.0, 0.0, 90.0)])
    B = hip.Mat(ctx, n, 2, hip.F64)
    hip.rhs_fill(ctx, v["mesh"], B, hip.CFIE, k, ALPHA, inc)
    X = hip.Mat(ctx, n, 2, hip.F64)
    Xg = hip.Mat(ctx, n, 2, hip.F64)
    rep, ok = v["op"].bicgstab(B, X, tol=1e-6, max_iter=500)
    assert ok and rep.relres < 1e-5
    rep2, ok2 = v["op"].gmres(B, Xg, m=30, tol=1e-6, max_restart=20)
    assert ok2
    assert rel(X.numpy(), Xg.numpy()) < 1e-3
    for m in (B, X, Xg):
        m.close()


def test_sibling_grouped_translation_gives_the_same_matvec(ctx, built, monkeypatch):
    # the opt-in kernel visits each child's far boxes in far-list order: same sums
    v = built["s4_800MHz"]
    x = _vec(v["rwg"].shape[0], 9)
    y0 = _apply(ctx, v["op"], x)
    monkeypatch.setenv("EMFA_MLFMA_SIBLING", "1")
    y1 = _apply(ctx, v["op"], x)
    assert np.array_equal(y0, y1)


def test_multi_item_loops_give_the_same_matvec(ctx, built, monkeypatch):
    # aggregation, upward pass and translation take several unknowns /
    # children / far boxes per loop step with their loads issued together,
    # summed in list order: the bits of the one-item loops
    for name in CASES:
        v = built[name]
        x = _vec(v["rwg"].shape[0], 10)
        y4 = _apply(ctx, v["op"], x)
        monkeypatch.setenv("EMFA_MLFMA_UNROLL", "1")
        y1 = _apply(ctx, v["op"], x)
        monkeypatch.delenv("EMFA_MLFMA_UNROLL")
        assert np.array_equal(y4, y1), name


def test_preconditioned_operator_matches_oracle_and_speeds_bicgstab(ctx):
    """MLFMA::calculatePreconditioner (MLFMA.cpp:1294-1391) on the device:
    P's nonzero count equals the restatement's, the preconditioned matvec
    P Z x matches P_oracle (near + far) x, and BiCGSTAB on P Z X = P B
    reaches the unpreconditioned solution.  With the shipped thresholds the
    reference's preconditioner (a right approximate inverse applied on the
    left, MLFMA.cpp:406-407,506-508) leaves the step count of this combined
    operator unchanged (20 and 20 to 1e-6); the CPU restatement shows it
    slowing the electric-field operator, a property of the algorithm."""
    xyz, tri, rwg = sphere(3, 0.5)
    f = 1.2e9
    k = hip.wavenumber(f)
    mesh = hip.Mesh(ctx, xyz, tri, rwg)
    op = hip.MLFMA(ctx, mesh, k, ALPHA, 0.25)
    plain = hip.MLFMA(ctx, mesh, k, ALPHA, 0.25)
    T = OM.Tree(xyz, tri, rwg, f, 0.25)
    n = rwg.shape[0]
    nnz = op.precondition(mesh, 0.9, 0.87)
    Z = hip.Mat(ctx, n, n, hip.F64)
    hip.fill(ctx, mesh, Z, hip.CFIE, k, 3, alpha=ALPHA)
    z = Z.numpy()
    Z.close()
    rows, cols = T.near_pairs()
    Zn = np.zeros_like(z)
    Zn[rows, cols] = z[rows, cols]
    mid = (xyz[rwg[:, 0]] + xyz[rwg[:, 1]]) / 2
    P = OM.sai(T.boxes[T.level], mid, Zn, 0.9, 0.87, T.blen[T.level])
    assert nnz == np.count_nonzero(P)
    x = _vec(n, 21)
    rad, rec = T.rad_rec(xyz, tri, rwg, ALPHA)
    y_ref = P @ (Zn @ x + T.far_matvec(x, rad, rec))
    # P is built from the FP32 near entries on the device: FP32 rounding times its conditioning
    assert rel(_apply(ctx, op, x), y_ref) < 1e-4
    inc = hip.inc_table([(0.0, 0.0, 0.0)])
    B = hip.Mat(ctx, n, 1, hip.F64)
    hip.rhs_fill(ctx, mesh, B, hip.CFIE, k, ALPHA, inc)
    X1, X0 = hip.Mat(ctx, n, 1, hip.F64), hip.Mat(ctx, n, 1, hip.F64)
    rep1, ok1 = op.bicgstab(B, X1, tol=1e-6, max_iter=500)
    rep0, ok0 = plain.bicgstab(B, X0, tol=1e-6, max_iter=500)
    print(f"BiCGSTAB steps: preconditioned {rep1.iterations}, plain {rep0.iterations}", flush=True)
    assert ok1 and ok0 and rep1.iterations <= rep0.iterations
    assert rel(X1.numpy(), X0.numpy()) < 1e-4
    for m in (B, X1, X0, op, plain, mesh):
        m.close()
