"""IP-SIE operator on the device (emfa_fastwt_*) against the numpy
restatement (oracle/fastwt.py), which tests/test_cpu_fastwt.py pins: the
matvec on conformal, coincident-vertex and nonconformal meshes (near entries,
four-channel patterns, the MLFMA passes), the excitation and the RCS; on a
conformal mesh the half-RWG operator recombines into the RWG MLFMA operator;
BiCGSTAB on it solves the restated system, and with the sparse approximate
inverse reaches the same answer."""
import sys

import numpy as np
import pytest

import meshgen
from conftest import rel, sphere
from emfa_amd import hip
from oracle import fastwt as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu

ALPHA = 0.5
BETA = W.beta(0.1, 0.07503)
# name: (mesh builder, frequency, box threshold)
CASES = {
    "sphere_s3": (lambda: hip.icosphere(3, 0.5), 1.2e9, 0.25),
    "cube_coincident": (lambda: meshgen.cube_faces([5] * 6, 1.0), 6e8, 0.2),
    "cube_nonconformal": (lambda: meshgen.cube_faces([4, 5, 4, 5, 4, 5], 1.0), 6e8, 0.2),
    "plate_pair": (lambda: meshgen.plate_pair(9, 12, 1.0), 9e8, 0.25),
}


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _vec(n, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(n) + 1j * rng.standard_normal(n)


def _contour(xyz, tri):
    c = hip.contour_build(xyz, tri)
    return c, W.Contour(c.hrwg, c.length, c.line_ptr, c.line_tri, c.line_seg)


@pytest.fixture(scope="module")
def built(ctx):
    out = {}
    for name, (make, f, thr) in CASES.items():
        xyz, tri = make()
        rwg, _ = hip.rwg_build(tri)
        c, C = _contour(xyz, tri)
        k = hip.wavenumber(f)
        mesh = hip.Mesh(ctx, xyz, tri, rwg)
        _say(f"fastwt {name}: {C.n} half-RWGs, {(C.hrwg[:, 5] == 1).sum()} nonconformal")
        op = hip.FastWT(ctx, mesh, c, k, ALPHA, thr, BETA)
        T = W.tree(xyz, tri, C, f, thr)
        out[name] = dict(xyz=xyz, tri=tri, rwg=rwg, c=c, C=C, k=k, f=f, mesh=mesh, op=op, T=T)
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


def _oracle_op(v):
    """(near matrix on the near pattern, rad, rec) of the restatement."""
    if "near" not in v:
        C, T = v["C"], v["T"]
        v["near"] = W.near_entries(v["xyz"], v["tri"], C, T, v["k"], ALPHA, BETA)
        v["rad"], v["rec"] = W.patterns(T, v["xyz"], v["tri"], C, ALPHA)
    return v["near"], v["rad"], v["rec"]


@pytest.mark.parametrize("name", list(CASES))
def test_tree_matches_oracle(built, name):
    v = built[name]
    info, T = v["op"].info(), v["T"]
    assert info["levels"] == T.level
    assert info["boxes"][1:] == [len(b) for b in T.boxes[1:]]
    assert info["multipoles"][2:] == [T.L[l] for l in range(2, T.level + 1)]
    Zn, _, _ = _oracle_op(v)
    assert info["near_nnz"] == int(np.count_nonzero(W.near_mask(T, v["C"].n)))


@pytest.mark.parametrize("name", list(CASES))
def test_matvec_matches_oracle(ctx, built, name):
    v = built[name]
    _say(f"fastwt matvec {name}")
    Zn, rad, rec = _oracle_op(v)
    n = v["C"].n
    for seed in (1, 2):
        x = _vec(n, seed)
        y = _apply(ctx, v["op"], x)
        yo = W.matvec(v["T"], Zn, rad, rec, x)
        assert rel(y, yo) < 3e-5, seed
        # the near part alone and the far part alone are both visible at this size
        assert rel(Zn @ x, yo) > 1e-2 and rel(yo - Zn @ x, yo) > 1e-3


def test_matvec_is_repeatable(ctx, built):
    v = built["cube_nonconformal"]
    x = _vec(v["C"].n, 7)
    assert np.array_equal(_apply(ctx, v["op"], x), _apply(ctx, v["op"], x))


def test_multi_item_loops_give_the_same_matvec(ctx, built, monkeypatch):
    # four channels: several unknowns / children / far boxes per loop step
    v = built["cube_nonconformal"]
    x = _vec(v["C"].n, 8)
    y4 = _apply(ctx, v["op"], x)
    monkeypatch.setenv("EMFA_MLFMA_UNROLL", "1")
    y1 = _apply(ctx, v["op"], x)
    assert np.array_equal(y4, y1)


def test_conformal_halves_recombine_into_the_rwg_operator(ctx, built):
    """On the conformal sphere the half-RWG tree equals the RWG tree and the
    line and penalty parts cancel: P^T A_wt P = A_rwg (FP32 operators)."""
    v = built["sphere_s3"]
    P = W.expand(v["C"], v["rwg"])
    op_rwg = hip.MLFMA(ctx, v["mesh"], v["k"], ALPHA, CASES["sphere_s3"][2])
    try:
        for seed in (3, 4):
            x = _vec(v["rwg"].shape[0], seed)
            y = P.T @ _apply(ctx, v["op"], P @ x)
            yr = _apply(ctx, op_rwg, x)
            assert rel(y, yr) < 2e-5, seed
    finally:
        op_rwg.close()


@pytest.mark.parametrize("name", ["cube_nonconformal", "plate_pair"])
def test_excitation_and_rcs_match_oracle(ctx, built, name):
    v = built[name]
    n, k = v["C"].n, v["k"]
    inc = O.inc_table([(0.0, 0.0, 0.0), (60.0, 30.0, 90.0), (120.0, 250.0, 45.0)])
    B = hip.Mat(ctx, n, 3, hip.F64)
    v["op"].rhs(B, inc)
    Bo = W.rhs(v["xyz"], v["tri"], v["C"], k, ALPHA, inc)
    assert rel(B.numpy(), Bo) < 1e-12
    th = np.radians(np.arange(0, 361, 15.0))
    dirs = np.stack([np.sin(th), 0 * th, np.cos(th)], 1)
    X = hip.Mat.from_numpy(ctx, np.stack([_vec(n, 5), _vec(n, 6)], 1), hip.F64)
    for col in (0, 1):
        r = v["op"].rcs(X, dirs, col)
        ro = W.rcs(v["xyz"], v["tri"], v["C"], k, X.numpy()[:, col], dirs)
        assert np.abs(r - ro).max() < 1e-9
    B.close()
    X.close()


@pytest.mark.parametrize("name", ["cube_nonconformal", "plate_pair"])
def test_bicgstab_solves_the_restated_system(ctx, built, name):
    """BiCGSTAB on the operator (variant 0, the algorithm of
    FastWT::iterativeSolve): its answer leaves a small residual in the
    restated operator.  (Against the
    dense all-pairs IP-SIE solve the answers differ by the reference's own
    far-field error, up to 4 dB here with 0.2-wavelength boxes: the
    translation stops at l <= k D, D the box separation; the restatement's
    own solve of its operator shows the same difference.)"""
    v = built[name]
    n = v["C"].n
    inc = O.inc_table([(30.0, 20.0, 0.0)])
    B = hip.Mat(ctx, n, 1, hip.F64)
    v["op"].rhs(B, inc)
    X = hip.Mat(ctx, n, 1, hip.F64)
    rep, ok = v["op"].bicgstab(B, X, tol=1e-6, max_iter=2000, variant=0)
    _say(f"fastwt {name}: BiCGSTAB {rep.iterations} steps, residual {rep.relres:.2e}")
    assert ok and rep.relres < 1e-6
    Zn, rad, rec = _oracle_op(v)
    b = B.numpy()[:, 0]
    res = W.matvec(v["T"], Zn, rad, rec, X.numpy()[:, 0]) - b
    assert np.linalg.norm(res) / np.linalg.norm(b) < 1e-4
    B.close()
    X.close()


def test_preconditioned_operator_solves_the_same_system(ctx, built):
    """Sparse approximate inverse on the IP-SIE operator (the MLFMA's, over
    half-RWG midpoints): P Z x = P b reaches the unpreconditioned solution."""
    make, f, thr = CASES["cube_nonconformal"]
    v = built["cube_nonconformal"]
    n = v["C"].n
    op2 = hip.FastWT(ctx, v["mesh"], v["c"], v["k"], ALPHA, thr, BETA)
    try:
        nnz = op2.precondition(v["mesh"], 0.9, 0.87)
        assert nnz > n
        inc = O.inc_table([(0.0, 0.0, 0.0)])
        B = hip.Mat(ctx, n, 1, hip.F64)
        v["op"].rhs(B, inc)
        X1, X2 = hip.Mat(ctx, n, 1, hip.F64), hip.Mat(ctx, n, 1, hip.F64)
        r1, ok1 = v["op"].bicgstab(B, X1, tol=1e-7, max_iter=3000, variant=0)
        r2, ok2 = op2.bicgstab(B, X2, tol=1e-7, max_iter=3000, variant=0)
        _say(f"fastwt SAI: {r1.iterations} -> {r2.iterations} steps")
        assert ok1 and ok2
        assert rel(X2.numpy(), X1.numpy()) < 1e-4
        for m in (B, X1, X2):
            m.close()
    finally:
        op2.close()


def test_rejects_a_half_rwg_off_its_triangle(ctx, built):
    v = built["plate_pair"]
    c = v["c"]
    bad = hip.Contour(c.hrwg.copy(), c.length, c.line_ptr, c.line_tri, c.line_seg)
    t = bad.hrwg[0, 3]
    other = [x for x in range(len(v["tri"])) if x != t][0]
    bad.hrwg[0, 3] = other
    with pytest.raises(hip.EmfaError):
        hip.FastWT(ctx, v["mesh"], bad, v["k"], ALPHA, 0.25, BETA)
