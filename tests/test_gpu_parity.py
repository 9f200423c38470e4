"""Parity of the gfx950 path (through the C ABI) against the FP64 oracle.

Tolerances: the FP64 path must match the oracle to rounding (entry-wise
relative 1e-10 on Z, 1e-9 on solutions); the FP32 fill is held to 2e-5
relative (single-precision arithmetic on the same formulas); RCS in dB to
0.05 dB (FP64) and 0.1 dB (FP32 + refinement), the north-star bounds.
"""
import sys

import numpy as np
import pytest
import scipy.linalg as sl

from conftest import rel, sphere
from emfa_amd import hip
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _cases_dirs(th):
    th = np.radians(np.asarray(th, dtype=float))
    return np.stack([np.sin(th), 0 * th, np.cos(th)], 1)


# ---------------------------------------------------------------- products
@pytest.mark.parametrize("aprec", [hip.F64, hip.F32])
@pytest.mark.parametrize("shape", [(64, 48, 3), (300, 257, 17), (129, 1000, 181), (2500, 300, 40), (2100, 64, 20)])
def test_gemm_residual(ctx, aprec, shape):
    n, k, m = shape
    _say(f"gemm_residual aprec={aprec} shape={shape}")
    rng = np.random.default_rng(1)
    A = rng.standard_normal((n, k)) + 1j * rng.standard_normal((n, k))
    X = rng.standard_normal((k, m)) + 1j * rng.standard_normal((k, m))
    C0 = rng.standard_normal((n, m)) + 1j * rng.standard_normal((n, m))
    if aprec == hip.F32:
        A = A.astype(np.complex64).astype(np.complex128)
    dA = hip.Mat.from_numpy(ctx, A, aprec)
    dX = hip.Mat.from_numpy(ctx, X, hip.F64)
    dC = hip.Mat.from_numpy(ctx, C0, hip.F64)
    hip.gemm_residual(ctx, dA, dX, dC, negate=True, accumulate=True)
    ref = C0 - A @ X
    assert rel(dC.numpy(), ref) < 1e-13
    hip.gemm_residual(ctx, dA, dX, dC, negate=False, accumulate=False)
    assert rel(dC.numpy(), A @ X) < 1e-13


# ---------------------------------------------------------------- fill
@pytest.mark.parametrize("quad", [3, 7])
@pytest.mark.parametrize("subdiv", [1, 2])
def test_fill_efie_f64(ctx, quad, subdiv):
    _say(f"fill efie f64 quad={quad} subdiv={subdiv}")
    xyz, tri, rwg = sphere(subdiv, 0.5)
    k = hip.wavenumber(3e8)
    Zo = O.fill_efie(xyz, tri, rwg, k, quad)
    M = hip.Mesh(ctx, xyz, tri, rwg)
    Z = hip.Mat(ctx, M.n, M.n, hip.F64)
    hip.fill(ctx, M, Z, hip.EFIE, k, quad)
    Zg = Z.numpy()
    err = np.abs(Zg - Zo) / np.abs(Zo).max()
    assert err.max() < 1e-12, err.max()


def test_fill_efie_sigma_f64(ctx):
    _say("fill efie f64 sigma")
    xyz, tri, rwg = sphere(1, 0.5)
    k = hip.wavenumber(3e8)
    sig = 0.0025 + 0.001j
    Zo = O.fill_efie(xyz, tri, rwg, k, 3, sigma=sig)
    M = hip.Mesh(ctx, xyz, tri, rwg)
    Z = hip.Mat(ctx, M.n, M.n, hip.F64)
    hip.fill(ctx, M, Z, hip.EFIE, k, 3, sigma=sig)
    assert np.abs(Z.numpy() - Zo).max() / np.abs(Zo).max() < 1e-12


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("quad", [3, 7])
def test_fill_cfie_f64(ctx, mode, quad):
    _say(f"fill cfie f64 mode={mode} quad={quad}")
    xyz, tri, rwg = sphere(2, 1.0)
    k = hip.wavenumber(6e8)
    Zo = O.fill_cfie(xyz, tri, rwg, k, 0.5, quad, mode)
    M = hip.Mesh(ctx, xyz, tri, rwg)
    Z = hip.Mat(ctx, M.n, M.n, hip.F64)
    hip.fill(ctx, M, Z, hip.CFIE, k, quad, alpha=0.5, cfie_mode=mode)
    assert np.abs(Z.numpy() - Zo).max() / np.abs(Zo).max() < 1e-12


@pytest.mark.parametrize("kind", [hip.EFIE, hip.CFIE])
def test_fill_f32(ctx, kind):
    _say(f"fill f32 kind={kind}")
    xyz, tri, rwg = sphere(2, 1.0)
    k = hip.wavenumber(6e8)
    if kind == hip.EFIE:
        Zo = O.fill_efie(xyz, tri, rwg, k, 3)
    else:
        Zo = O.fill_cfie(xyz, tri, rwg, k, 0.5, 3, 0)
    M = hip.Mesh(ctx, xyz, tri, rwg)
    Z = hip.Mat(ctx, M.n, M.n, hip.F32)
    hip.fill(ctx, M, Z, kind, k, 3, alpha=0.5)
    assert np.abs(Z.numpy() - Zo).max() / np.abs(Zo).max() < 2e-5


def test_fill_chunked_equals_single(ctx, monkeypatch):
    """A small scratch budget (many test-triangle chunks) gives bit-identical Z."""
    _say("fill chunked")
    xyz, tri, rwg = sphere(2, 1.0)
    k = hip.wavenumber(6e8)
    M = hip.Mesh(ctx, xyz, tri, rwg)
    Z1 = hip.Mat(ctx, M.n, M.n, hip.F64)
    hip.fill(ctx, M, Z1, hip.CFIE, k, 3, alpha=0.5)
    monkeypatch.setenv("EMFA_FILL_SCRATCH_MB", "1")
    Z2 = hip.Mat(ctx, M.n, M.n, hip.F64)
    hip.fill(ctx, M, Z2, hip.CFIE, k, 3, alpha=0.5)
    assert np.array_equal(Z1.numpy(), Z2.numpy())


# ---------------------------------------------------------------- rhs / rcs
@pytest.mark.parametrize("kind", [hip.EFIE, hip.CFIE])
def test_rhs(ctx, kind):
    _say(f"rhs kind={kind}")
    xyz, tri, rwg = sphere(2, 1.0)
    k = hip.wavenumber(6e8)
    angles = [(th, ph, po) for th in (0, 37, 90, 180) for ph in (0, 45) for po in (0, 90)]
    inc = hip.inc_table(angles)
    Bo = O.rhs(xyz, tri, rwg, kind, k, 0.5, inc)
    M = hip.Mesh(ctx, xyz, tri, rwg)
    B = hip.Mat(ctx, M.n, len(angles), hip.F64)
    hip.rhs_fill(ctx, M, B, kind, k, 0.5, inc)
    assert rel(B.numpy(), Bo) < 1e-13


def test_rcs(ctx):
    _say("rcs")
    xyz, tri, rwg = sphere(2, 1.0)
    k = hip.wavenumber(6e8)
    rng = np.random.default_rng(3)
    X = rng.standard_normal((rwg.shape[0], 3)) + 1j * rng.standard_normal((rwg.shape[0], 3))
    dirs = _cases_dirs(np.arange(0, 361, 7.0))
    cols = np.arange(dirs.shape[0]) % 3
    ro = O.rcs(xyz, tri, rwg, k, X, dirs, cols)
    M = hip.Mesh(ctx, xyz, tri, rwg)
    dX = hip.Mat.from_numpy(ctx, X, hip.F64)
    rg = hip.rcs(ctx, M, k, dX, dirs, cols)
    assert np.abs(rg - ro).max() < 1e-9


# ---------------------------------------------------------------- LU
@pytest.mark.parametrize("n", [5, 64, 333, 700])
def test_lu_f64_matches_getrf(ctx, n):
    _say(f"lu f64 n={n}")
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    dA = hip.Mat.from_numpy(ctx, A, hip.F64)
    lu = hip.LU(ctx, dA, hip.F64)
    LUo, pivo, info = O.lu(A)
    assert info == 0 and lu.info() == 0
    assert np.array_equal(lu.pivots(), pivo)
    assert np.abs(lu.factor().numpy() - LUo).max() / np.abs(LUo).max() < 1e-10
    B = rng.standard_normal((n, 7)) + 1j * rng.standard_normal((n, 7))
    dB = hip.Mat.from_numpy(ctx, B, hip.F64)
    lu.solve(dB)
    assert rel(dB.numpy(), np.linalg.solve(A, B)) < 1e-10


@pytest.mark.parametrize("xchg", [0, 1])
def test_lu_f64_normal_launch_panels_match_getrf(ctx, monkeypatch, xchg):
    """The panel launched without the cooperative API (what a context does
    when the runtime refuses a cooperative launch, and what contexts sharing
    the device always do), for both pivot-exchange designs: getrf pivots."""
    _say(f"lu f64 normal-launch panel xchg={xchg}")
    monkeypatch.setenv("EMFA_PANEL_COOP", "0")
    monkeypatch.setenv("EMFA_PANEL_X", str(xchg))
    n = 777
    rng = np.random.default_rng(77 + xchg)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    dA = hip.Mat.from_numpy(ctx, A, hip.F64)
    lu = hip.LU(ctx, dA, hip.F64)
    LUo, pivo, info = O.lu(A)
    assert info == 0 and lu.info() == 0
    assert np.array_equal(lu.pivots(), pivo)
    assert np.abs(lu.factor().numpy() - LUo).max() / np.abs(LUo).max() < 1e-10
    B = rng.standard_normal((n, 3)) + 1j * rng.standard_normal((n, 3))
    dB = hip.Mat.from_numpy(ctx, B, hip.F64)
    lu.solve(dB)
    assert rel(dB.numpy(), np.linalg.solve(A, B)) < 1e-10
    for m in (lu, dA, dB):
        m.close()


@pytest.mark.parametrize("n", [100, 517])
def test_lu_f32(ctx, n):
    _say(f"lu f32 n={n}")
    rng = np.random.default_rng(n + 1)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) + 4 * np.sqrt(n) * np.eye(n)
    A32 = A.astype(np.complex64).astype(np.complex128)
    dA = hip.Mat.from_numpy(ctx, A32, hip.F32)
    lu = hip.LU(ctx, dA, hip.F32)
    B = rng.standard_normal((n, 5)) + 1j * rng.standard_normal((n, 5))
    dB = hip.Mat.from_numpy(ctx, B, hip.F32)
    lu.solve(dB)
    ref = np.linalg.solve(A32, B.astype(np.complex64).astype(np.complex128))
    assert rel(dB.numpy(), ref) < 1e-4


@pytest.mark.parametrize("n,m", [(70, 3), (333, 64), (700, 90)])
def test_lookback_solve_matches_blocked_solve(ctx, monkeypatch, n, m):
    # the opt-in one-launch triangular solve (look-back flags) against the
    # default blocked solve and LAPACK, FP64 factor
    _say(f"lookback solve n={n} m={m}")
    rng = np.random.default_rng(n + m)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    B = rng.standard_normal((n, m)) + 1j * rng.standard_normal((n, m))
    dA = hip.Mat.from_numpy(ctx, A, hip.F64)
    lu = hip.LU(ctx, dA, hip.F64)
    d0 = hip.Mat.from_numpy(ctx, B, hip.F64)
    lu.solve(d0)
    monkeypatch.setenv("EMFA_SOLVE_LB_MAX", "128")
    d1 = hip.Mat.from_numpy(ctx, B, hip.F64)
    lu.solve(d1)
    ref = np.linalg.solve(A, B)
    assert rel(d1.numpy(), ref) < 1e-10
    assert rel(d1.numpy(), d0.numpy()) < 1e-11


def test_autotuner_and_every_lu_parameter_set_keep_getrf_pivots(ctx):
    _say("autotune")
    n = 600
    rep = ctx.autotune_lu(n, hip.F64)
    assert rep["candidates"] and rep["nb"] in (256, 384, 512) and rep["best_ms"] > 0
    rep32 = ctx.autotune_lu(n, hip.F32, for_refinement=True)
    assert len(rep32["candidates"]) == 12 and rep32["update"] in ("fp32", "bf16x3")
    ctx.lu_update_mode(0)
    ctx.lu_params(0, 0, -1)
    rng = np.random.default_rng(77)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    LUo, pivo, _ = O.lu(A)
    for nb, g, x in [(256, 0, 1), (384, 0, 0), (512, 16, 1), (0, 0, -1)]:
        ctx.lu_params(nb, g, x)
        dA = hip.Mat.from_numpy(ctx, A, hip.F64)
        lu = hip.LU(ctx, dA, hip.F64)
        assert np.array_equal(lu.pivots(), pivo), (nb, g, x)
        assert np.abs(lu.factor().numpy() - LUo).max() / np.abs(LUo).max() < 1e-10
        lu.close()
        dA.close()
    ctx.lu_params(0, 0, -1)


def test_bf16x3_update_lu_refines_to_fp64_accuracy(ctx):
    # FP32 factor whose trailing updates run on bf16 MFMA (three-term split):
    # a few refinement steps reach the same 1e-12 residual
    _say("bf16x3 lu")
    xyz, tri, rwg = sphere(3, 1.0)
    k = hip.wavenumber(6e8)
    M = hip.Mesh(ctx, xyz, tri, rwg)
    Z = hip.Mat(ctx, M.n, M.n, hip.F32)
    hip.fill(ctx, M, Z, hip.CFIE, k, 3, alpha=0.5)
    inc = hip.inc_table([(t, 0, 0) for t in range(0, 181, 45)])
    B = hip.Mat(ctx, M.n, inc.shape[0], hip.F64)
    hip.rhs_fill(ctx, M, B, hip.CFIE, k, 0.5, inc)
    ref = np.linalg.solve(Z.numpy(), B.numpy())
    ctx.lu_update_mode(1)
    try:
        lu = hip.LU(ctx, Z, hip.F32)
    finally:
        ctx.lu_update_mode(0)
    X = hip.Mat(ctx, M.n, inc.shape[0], hip.F64)
    rep, ok = hip.refine(ctx, Z, lu, B, X, tol=1e-12, maxit=12)
    assert 1e-8 < rep.relres_first < 1e-2  # the split products keep ~16 bits
    assert ok and rep.relres_final < 1e-12
    assert rel(X.numpy(), ref) < 1e-10
    # the factor's pivots are a valid partial-pivoting sequence of the same matrix
    assert len(set(lu.pivots().tolist())) > 0


@pytest.mark.parametrize("n,nb", [(1003, 256), (700, 512), (1600, 384)])
def test_bf16x3_update_ragged_sizes(ctx, n, nb):
    # ragged outer blocks (K not a multiple of 16, M not of 256) go through the
    # zero-padded operand planes; the factor matches the FP32-MFMA one and
    # refines to the FP64 solution
    _say(f"bf16x3 ragged n={n} nb={nb}")
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) + 4 * np.sqrt(n) * np.eye(n)
    B = rng.standard_normal((n, 3)) + 1j * rng.standard_normal((n, 3))
    Z = hip.Mat.from_numpy(ctx, A, hip.F32)
    ctx.lu_params(nb, 0, -1)
    try:
        lu0 = hip.LU(ctx, Z, hip.F32)
        ctx.lu_update_mode(1)
        lu1 = hip.LU(ctx, Z, hip.F32)
        ctx.lu_params(nb, 0, 2)  # and without interchanges
        lu2 = hip.LU(ctx, Z, hip.F32)
    finally:
        ctx.lu_update_mode(0)
        ctx.lu_params(0, 0, -1)
    # diagonally dominant: partial pivoting keeps the diagonal, so all three agree
    assert np.array_equal(lu0.pivots(), np.arange(n)) and np.array_equal(lu1.pivots(), np.arange(n))
    assert lu2.pivoting() == (0, 0) and np.array_equal(lu2.pivots(), np.arange(n))
    f0 = lu0.factor().numpy()
    for lu in (lu1, lu2):
        assert np.abs(lu.factor().numpy() - f0).max() / np.abs(f0).max() < 1e-5
    Bd = hip.Mat.from_numpy(ctx, B, hip.F64)
    X = hip.Mat(ctx, n, 3, hip.F64)
    for lu in (lu1, lu2):
        rep, ok = hip.refine(ctx, Z, lu, Bd, X, tol=1e-12, maxit=12)
        assert ok and rep.relres_final < 1e-12
        assert rel(X.numpy(), np.linalg.solve(Z.numpy(), B)) < 1e-10
    for m in (lu0, lu1, lu2, Z, Bd, X):
        m.close()


def test_factor_without_interchanges_on_cfie_and_its_growth_fallback(ctx):
    _say("no-interchange LU")
    xyz, tri, rwg = sphere(3, 1.0)
    k = hip.wavenumber(6e8)
    M = hip.Mesh(ctx, xyz, tri, rwg)
    Z = hip.Mat(ctx, M.n, M.n, hip.F32)
    hip.fill(ctx, M, Z, hip.CFIE, k, 3, alpha=0.5)
    inc = hip.inc_table([(t, 0, 0) for t in range(0, 181, 60)])
    B = hip.Mat(ctx, M.n, inc.shape[0], hip.F64)
    hip.rhs_fill(ctx, M, B, hip.CFIE, k, 0.5, inc)
    X = hip.Mat(ctx, M.n, inc.shape[0], hip.F64)
    ctx.lu_params(256, 0, 2)
    try:
        lu = hip.LU(ctx, Z, hip.F32)
        # the CFIE matrix factors without interchanges and refines to FP64 accuracy
        assert lu.pivoting() == (0, 0)
        rep, ok = hip.refine(ctx, Z, lu, B, X, tol=1e-12, maxit=12)
        assert ok and rep.relres_final < 1e-12 and rep.relres_first < 1e-3
        assert rel(X.numpy(), np.linalg.solve(Z.numpy(), B.numpy())) < 1e-10
        lu.close()
        # a tiny leading pivot makes the multipliers of column 0 ~1e4: redone with pivoting
        n = 700
        rng = np.random.default_rng(3)
        A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) + 4 * np.sqrt(n) * np.eye(n)
        A[0, 0] = 1e-4
        Zs = hip.Mat.from_numpy(ctx, A, hip.F32)
        lus = hip.LU(ctx, Zs, hip.F32)
        assert lus.pivoting() == (1, 1)
        ctx.lu_params(256, 0, 1)
        lup = hip.LU(ctx, Zs, hip.F32)
        assert np.array_equal(lus.pivots(), lup.pivots()) and lus.pivots()[0] != 0
        assert np.array_equal(lus.factor().numpy(), lup.factor().numpy())
        # FP64 factorisations keep their interchanges whatever the variant
        ctx.lu_params(256, 0, 2)
        lud = hip.LU(ctx, hip.Mat.from_numpy(ctx, A, hip.F64), hip.F64)
        assert lud.pivoting() == (1, 0)
        LUo, pivo, _ = O.lu(A)
        assert np.array_equal(lud.pivots(), pivo)
        for m in (lus, lup, lud, Zs):
            m.close()
    finally:
        ctx.lu_params(0, 0, -1)
    for m in (Z, B, X):
        m.close()
    M.close()


@pytest.mark.parametrize("prec", [hip.F32, hip.F64])
@pytest.mark.parametrize("n,m", [(700, 1), (1003, 7), (513, 70), (256, 3), (2600, 40), (2600, 30)])
def test_inverse_block_solve_matches_blocked_substitution(ctx, monkeypatch, prec, n, m):
    # the solve through the diagonal-block inverses against the blocked
    # substitution (no inverses) on the same factor data, pivoted and not;
    # n = 2600 takes the tall skinny updates (>= 2048 rows)
    _say(f"dinv solve n={n} m={m} prec={prec}")
    rng = np.random.default_rng(n + m)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    if n > 2000:
        A += 4 * np.sqrt(n) * np.eye(n)  # keeps the FP32 solution error well inside the tolerance
    B = rng.standard_normal((n, m)) + 1j * rng.standard_normal((n, m))
    Z = hip.Mat.from_numpy(ctx, A, prec)
    lu = hip.LU(ctx, Z, prec)
    d1 = hip.Mat.from_numpy(ctx, B, prec)
    lu.solve(d1)
    monkeypatch.setenv("EMFA_SOLVE_BLOCKED", "1")
    d0 = hip.Mat.from_numpy(ctx, B, prec)
    lu.solve(d0)
    # unshifted Gaussian matrices (kappa ~ n) pivot; FP32 solutions carry kappa * 6e-8
    tol = 2e-3 if prec == hip.F32 else 1e-10
    assert rel(d1.numpy(), d0.numpy()) < tol
    assert rel(d1.numpy(), np.linalg.solve(Z.numpy(), B)) < tol
    for x in (lu, Z, d0, d1):
        x.close()


@pytest.mark.parametrize("n,m", [(1100, 65), (2600, 181)])
def test_mfma_block_apply_matches_valu_apply(ctx, monkeypatch, n, m):
    # more than 64 FP32 right-hand sides: the diagonal-block applies run as
    # MFMA products of the expanded blocks; k_dinv_apply (EMFA_APPLY_MFMA=0)
    # gives the same solution to FP32 accuracy on the same factor
    _say(f"mfma block apply n={n} m={m}")
    rng = np.random.default_rng(n + 7 * m)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) + 4 * np.sqrt(n) * np.eye(n)
    B = rng.standard_normal((n, m)) + 1j * rng.standard_normal((n, m))
    Z = hip.Mat.from_numpy(ctx, A, hip.F32)
    lu = hip.LU(ctx, Z, hip.F32)
    d1 = hip.Mat.from_numpy(ctx, B, hip.F32)
    lu.solve(d1)
    monkeypatch.setenv("EMFA_APPLY_MFMA", "0")
    d0 = hip.Mat.from_numpy(ctx, B, hip.F32)
    lu.solve(d0)
    assert rel(d1.numpy(), d0.numpy()) < 1e-4
    assert rel(d1.numpy(), np.linalg.solve(A, B)) < 2e-3
    for x in (lu, Z, d0, d1):
        x.close()


@pytest.mark.parametrize("m", [40, 64, 20])
def test_shared_context_solve_matches_skinny_products(monkeypatch, m):
    # a context with a CU share (as the solve server's workers) solves 32-64
    # columns with the expanded-operand MFMA GEMM instead of k_smallmm; 20
    # columns stay on k_smallmm.  Same factor, same answer to FP32 accuracy.
    _say(f"shared-context solve m={m}")
    n = 2600
    rng = np.random.default_rng(77 + m)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) + 4 * np.sqrt(n) * np.eye(n)
    B = rng.standard_normal((n, m)) + 1j * rng.standard_normal((n, m))
    c = hip.Context(0)
    mats = []
    try:
        c.set_cu_share(32)
        Z = hip.Mat.from_numpy(c, A, hip.F32)
        lu = hip.LU(c, Z, hip.F32)
        d1 = hip.Mat.from_numpy(c, B, hip.F32)
        mats += [Z, lu, d1]
        lu.solve(d1)
        monkeypatch.setenv("EMFA_SKINNY_MAX", "64")
        d0 = hip.Mat.from_numpy(c, B, hip.F32)
        mats.append(d0)
        lu.solve(d0)
        assert rel(d1.numpy(), d0.numpy()) < 1e-4
        assert rel(d1.numpy(), np.linalg.solve(A, B)) < 2e-3
    finally:
        for x in mats:
            x.close()
        c.close()


def test_lu_singular_reports(ctx):
    _say("lu singular")
    n = 40
    rng = np.random.default_rng(5)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    A[:, 17] = 0
    dA = hip.Mat.from_numpy(ctx, A, hip.F64)
    lu = hip.LU(ctx, dA, hip.F64)
    assert lu.singular and lu.info() == 18


def test_refine_f32_to_f64(ctx):
    _say("refine")
    xyz, tri, rwg = sphere(2, 1.0)
    k = hip.wavenumber(6e8)
    M = hip.Mesh(ctx, xyz, tri, rwg)
    Z = hip.Mat(ctx, M.n, M.n, hip.F32)
    hip.fill(ctx, M, Z, hip.CFIE, k, 3, alpha=0.5)
    inc = hip.inc_table([(t, 0, 0) for t in range(0, 181, 30)])
    B = hip.Mat(ctx, M.n, inc.shape[0], hip.F64)
    hip.rhs_fill(ctx, M, B, hip.CFIE, k, 0.5, inc)
    lu = hip.LU(ctx, Z, hip.F32)
    X = hip.Mat(ctx, M.n, inc.shape[0], hip.F64)
    rep, ok = hip.refine(ctx, Z, lu, B, X, tol=1e-12, maxit=10)
    assert ok and rep.relres_final < 1e-12 and rep.relres_first > rep.relres_final
    Zh = Z.numpy()
    ref = np.linalg.solve(Zh, B.numpy())
    assert rel(X.numpy(), ref) < 1e-10


def test_gmres(ctx):
    _say("gmres")
    xyz, tri, rwg = sphere(2, 1.0)
    k = hip.wavenumber(6e8)
    M = hip.Mesh(ctx, xyz, tri, rwg)
    Z = hip.Mat(ctx, M.n, M.n, hip.F64)
    hip.fill(ctx, M, Z, hip.CFIE, k, 3, alpha=0.5)
    inc = hip.inc_table([(0, 0, 0), (90, 0, 90)])
    B = hip.Mat(ctx, M.n, 2, hip.F64)
    hip.rhs_fill(ctx, M, B, hip.CFIE, k, 0.5, inc)
    X = hip.Mat(ctx, M.n, 2, hip.F64)
    rep, ok = hip.gmres(ctx, Z, B, X, m=30, tol=1e-10, max_restart=50)
    assert ok, rep.relres
    ref = np.linalg.solve(Z.numpy(), B.numpy())
    assert rel(X.numpy(), ref) < 1e-8


@pytest.mark.parametrize("variant", [0, 1])
@pytest.mark.parametrize("prec", [hip.F64, hip.F32])
def test_bicgstab_matches_oracle(ctx, variant, prec):
    """Device BiCGSTAB against the oracle's restatement of the same overload
    (variant 0: EMIterative/iml.cpp:3-39 at CFIE.cpp:76's eps 0.01 / 500;
    variant 1: iml.cpp:88-154) on the same matrix: the same step count per
    right-hand side and the same iterate to 1e-9 (FP64 vectors and products
    on both sides; an FP32 matrix is widened exactly).  At eps 1e-10 both
    variants reach the LU solution."""
    _say(f"bicgstab variant={variant} prec={prec}")
    xyz, tri, rwg = sphere(2, 1.0)
    k = hip.wavenumber(6e8)
    M = hip.Mesh(ctx, xyz, tri, rwg)
    Z = hip.Mat(ctx, M.n, M.n, prec)
    hip.fill(ctx, M, Z, hip.CFIE, k, 3, alpha=0.5)
    inc = hip.inc_table([(0, 0, 0), (70, 20, 90), (180, 0, 0)])
    B = hip.Mat(ctx, M.n, inc.shape[0], hip.F64)
    hip.rhs_fill(ctx, M, B, hip.CFIE, k, 0.5, inc)
    X = hip.Mat(ctx, M.n, inc.shape[0], hip.F64)
    Zh, Bh = Z.numpy(), B.numpy()
    rep, ok = hip.bicgstab(ctx, Z, B, X, tol=1e-2, max_iter=500, variant=variant)
    assert ok
    steps = 0
    for j in range(inc.shape[0]):
        if variant == 0:
            xo, it, oko = O.bicgstab(Zh, Bh[:, j], 1e-2, 500)
        else:
            xo, it, _ = O.bicgstab_mf(Zh, Bh[:, j], 1e-2, 500)
        steps += it
        assert rel(X.numpy()[:, j], xo) < 1e-9
    assert rep.iterations == steps
    rep2, ok2 = hip.bicgstab(ctx, Z, B, X, tol=1e-10, max_iter=500, variant=variant)
    assert ok2 and rep2.relres < 1e-9
    assert rel(X.numpy(), np.linalg.solve(Zh, Bh)) < 1e-7
    # too few steps: variant 0 reports it as the reference's `false`
    rep3, ok3 = hip.bicgstab(ctx, Z, B, X, tol=1e-12, max_iter=2, variant=variant)
    assert ok3 == (variant == 1) and rep3.max_steps == 2
    for m in (X, B, Z, M):
        m.close()


# ---------------------------------------------------------------- end to end (C1 mesh)
@pytest.mark.parametrize("quad", [3, 7])
def test_end_to_end_efie_sphere(ctx, quad):
    """EFIE, PEC sphere r=0.5 m, 300 MHz, icosphere s=3 (1920 RWG), x-pol from +z,
    bistatic RCS at 361 angles.  quad=7 is the configured workload (7-point
    outer and inner rules); quad=3 is the reference's own fill rule.  Bounds:
    FP64 path surface-current coefficients rel L2 <= 1e-5 and RCS <= 0.05 dB,
    FP32 + FP64 refinement RCS <= 0.1 dB, both against the FP64 oracle."""
    _say(f"end-to-end C1 quad={quad}")
    xyz, tri, rwg = sphere(3, 0.5)
    k = hip.wavenumber(3e8)
    inc = hip.inc_table([(0, 0, 0)])
    Zo = O.fill_efie(xyz, tri, rwg, k, quad)
    Bo = O.rhs(xyz, tri, rwg, 0, k, 1.0, inc)
    Xo = sl.solve(Zo, Bo)
    th = np.arange(0, 361, 1.0)
    dirs = _cases_dirs(th)
    ro = O.rcs(xyz, tri, rwg, k, Xo, dirs)
    M = hip.Mesh(ctx, xyz, tri, rwg)
    # FP64 plain path
    Z = hip.Mat(ctx, M.n, M.n, hip.F64)
    hip.fill(ctx, M, Z, hip.EFIE, k, quad)
    B = hip.Mat(ctx, M.n, 1, hip.F64)
    hip.rhs_fill(ctx, M, B, hip.EFIE, k, 1.0, inc)
    lu = hip.LU(ctx, Z, hip.F64)
    lu.solve(B)
    X = B.numpy()
    assert rel(X, Xo) <= 1e-5
    mag, _ = O.current(xyz, tri.shape[0], rwg, X)
    mago, _ = O.current(xyz, tri.shape[0], rwg, Xo)
    assert rel(mag, mago) <= 1e-5
    r64 = hip.rcs(ctx, M, k, B, dirs)
    _say(f"  FP64 path: current rel {rel(X, Xo):.2e}, RCS max |diff| {np.abs(r64 - ro).max():.2e} dB")
    assert np.abs(r64 - ro).max() <= 0.05
    # FP32 fill + FP32 LU + FP64 refinement
    Z32 = hip.Mat(ctx, M.n, M.n, hip.F32)
    hip.fill(ctx, M, Z32, hip.EFIE, k, quad)
    B2 = hip.Mat(ctx, M.n, 1, hip.F64)
    hip.rhs_fill(ctx, M, B2, hip.EFIE, k, 1.0, inc)
    lu32 = hip.LU(ctx, Z32, hip.F32)
    X2 = hip.Mat(ctx, M.n, 1, hip.F64)
    rep, ok = hip.refine(ctx, Z32, lu32, B2, X2, tol=1e-10, maxit=10)
    assert ok
    r32 = hip.rcs(ctx, M, k, X2, dirs)
    _say(f"  FP32+refine: RCS max |diff| {np.abs(r32 - ro).max():.2e} dB")
    assert np.abs(r32 - ro).max() <= 0.1


def test_batched_factor_matches_single_factors_and_redoes_a_grown_member(ctx):
    _say("batched LU")
    ctx.lu_params(256, 0, 2)
    ctx.lu_update_mode(1)
    rng = np.random.default_rng(11)
    try:
        for sub in (2, 3):  # 480 and 1920 unknowns: a partial last block, several outer steps
            xyz, tri, rwg = sphere(sub, 1.0)
            M = hip.Mesh(ctx, xyz, tri, rwg)
            Zs = []
            for f in (2e8, 3e8, 4.5e8):
                Z = hip.Mat(ctx, M.n, M.n, hip.F32)
                hip.fill(ctx, M, Z, hip.CFIE, hip.wavenumber(f), 3, alpha=0.5)
                Zs.append(Z)
            single = [hip.LU(ctx, Z, hip.F32) for Z in Zs]
            batch = hip.lu_factor_batch(ctx, Zs, hip.F32)
            B0 = (rng.standard_normal((M.n, 5)) + 1j * rng.standard_normal((M.n, 5))).astype(np.complex64)
            for s1, b1 in zip(single, batch):
                assert b1.pivoting() == (0, 0) and s1.pivoting() == (0, 0)
                Fs, Fb = s1.factor().numpy(), b1.factor().numpy()
                _say(f"n={M.n} factor identical: {np.array_equal(Fs, Fb)}, rel {rel(Fb, Fs):.2e}")
                assert rel(Fb, Fs) < 1e-6
                Xs, Xb = hip.Mat.from_numpy(ctx, B0, hip.F32), hip.Mat.from_numpy(ctx, B0, hip.F32)
                s1.solve(Xs)
                b1.solve(Xb)
                assert rel(Xb.numpy(), Xs.numpy()) < 1e-5
                for m in (s1, b1, Xs, Xb):
                    m.close()
            for Z in Zs:
                Z.close()
            M.close()
        # a member whose multipliers grow is redone alone with interchanges
        n = 700
        A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) + 4 * np.sqrt(n) * np.eye(n)
        G = A.copy()
        G[0, 0] = 1e-4
        Zo, Zg, Zo2 = (hip.Mat.from_numpy(ctx, X, hip.F32) for X in (A, G, A.T.copy()))
        lb = hip.lu_factor_batch(ctx, [Zo, Zg, Zo2], hip.F32)
        assert [l.pivoting() for l in lb] == [(0, 0), (1, 1), (0, 0)]
        lg = hip.LU(ctx, Zg, hip.F32)  # alone: the same fallback
        assert lg.pivoting() == (1, 1)
        assert np.array_equal(lb[1].pivots(), lg.pivots())
        assert np.array_equal(lb[1].factor().numpy(), lg.factor().numpy())
        lo = hip.LU(ctx, Zo, hip.F32)
        assert rel(lb[0].factor().numpy(), lo.factor().numpy()) < 1e-6
        for m in lb + [lg, lo, Zo, Zg, Zo2]:
            m.close()
    finally:
        ctx.lu_params(0, 0, -1)
        ctx.lu_update_mode(0)


@pytest.mark.parametrize("variant", [1, 2])
def test_one_launch_solve_steps_match_two_launch_steps(ctx, monkeypatch, variant):
    # EMFA_SOLVE_GH=1: FP32 factors keep L21 DL^-1 / U12 DU^-1 and solve with
    # one launch per 256-row step; answers match the two-launch steps
    _say(f"scaled-block solve variant={variant}")
    rng = np.random.default_rng(17 + variant)
    n = 1300
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) + 3 * np.sqrt(n) * np.eye(n)
    ctx.lu_params(256, 0, variant)
    ctx.lu_update_mode(1)
    try:
        Z = hip.Mat.from_numpy(ctx, A, hip.F32)
        lu0 = hip.LU(ctx, Z, hip.F32)
        monkeypatch.setenv("EMFA_SOLVE_GH", "1")
        lu1 = hip.LU(ctx, Z, hip.F32)
        Zs = [hip.Mat.from_numpy(ctx, X, hip.F32) for X in (A, A.T.copy())]
        lub = hip.lu_factor_batch(ctx, Zs, hip.F32)
        for m in (1, 13, 40, 64):
            B = (rng.standard_normal((n, m)) + 1j * rng.standard_normal((n, m))).astype(np.complex64)
            x0, x1, xb = (hip.Mat.from_numpy(ctx, B, hip.F32) for _ in range(3))
            lu0.solve(x0)
            lu1.solve(x1)
            lub[0].solve(xb)
            ref = x0.numpy()
            assert rel(x1.numpy(), ref) < 1e-5 and rel(xb.numpy(), ref) < 1e-5
            assert rel(ref, np.linalg.solve(A.astype(np.complex64).astype(np.complex128), B)) < 1e-4
            for x in (x0, x1, xb):
                x.close()
        for m in [lu0, lu1, Z] + lub + Zs:
            m.close()
    finally:
        ctx.lu_params(0, 0, -1)
        ctx.lu_update_mode(0)
