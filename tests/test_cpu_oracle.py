"""Pins of the FP64 oracle (CPU only).

The reference ships no test, golden vector or fixture for this path, so the
oracle is pinned against what the mathematics fixes, each check chosen so a
plausible slip (dropped term, wrong sign, transposed operand, wrong index)
fails it:
  - triangle rules: exactness on monomials (degree 2 / degree 5);
  - RWG topology: brute-force edge enumeration and RWG invariants;
  - analytic self-term potentials: independent Duffy-transformed quadrature;
  - reciprocity: the EFIE Galerkin matrix is symmetric off the self terms;
  - LU: LAPACK getrf (scipy) pivots and factors; GMRES/BiCGSTAB vs LU;
  - the whole chain fill -> solve -> RCS against the Mie series of a PEC
    sphere (EFIE, CFIE, MFIE).
"""
import math

import numpy as np
import pytest
import scipy.integrate as si
import scipy.linalg as sl

import meshgen
import mie
from oracle import oracle as O


# ---------------------------------------------------------------- quadrature
@pytest.mark.parametrize("q,deg", [(3, 2), (7, 5)])
def test_quadrature_exactness(q, deg):
    bary, w = O.quad_rule(q)
    assert abs(w.sum() - 1.0) < 1e-13
    # reference triangle (0,0),(1,0),(0,1): integral of x^a y^b = a! b! / (a+b+2)!
    V = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]])
    P = bary @ V
    for a in range(deg + 1):
        for b in range(deg + 1 - a):
            exact = math.factorial(a) * math.factorial(b) / math.factorial(a + b + 2)
            approx = 0.5 * np.sum(w * P[:, 0] ** a * P[:, 1] ** b)
            assert abs(approx - exact) < 1e-12, (q, a, b)
    # and not exact one degree higher (the rule is what it claims, no more)
    a = deg + 1
    exact = math.factorial(a) / math.factorial(a + 2)
    assert abs(0.5 * np.sum(w * P[:, 0] ** a) - exact) > 1e-6


# ---------------------------------------------------------------- RWG topology
def _brute_force_shared_edges(tri):
    edges = {}
    for t, (a, b, c) in enumerate(tri):
        for u, v in ((a, b), (b, c), (c, a)):
            edges.setdefault(frozenset((u, v)), []).append(t)
    return {e: ts for e, ts in edges.items() if len(ts) == 2}


@pytest.mark.parametrize("mesh", ["ico1", "ico2", "plate3", "ellipsoid1"])
def test_rwg_invariants_and_brute_force(mesh):
    if mesh.startswith("ico"):
        xyz, tri = meshgen.icosphere(int(mesh[-1]), 1.0)
    elif mesh.startswith("plate"):
        xyz, tri = meshgen.plate(int(mesh[-1]))
    else:
        xyz, tri = meshgen.ellipsoid(1, 1.0, 0.5, 0.3)
    rwg, nb = O.build_rwg(tri)
    shared = _brute_force_shared_edges(tri)
    assert rwg.shape[0] == len(shared)
    seen = set()
    for v1, v2, vp, vm, tp, tm in rwg:
        e = frozenset((v1, v2))
        assert e in shared and sorted(shared[e]) == sorted([tp, tm])
        seen.add(e)
        # free vertices are the opposite corners; v1 -> v2 runs as in the plus triangle
        assert set(tri[tp]) == {v1, v2, vp} and set(tri[tm]) == {v1, v2, vm}
        i = list(tri[tp]).index(v1)
        assert tri[tp][(i + 1) % 3] == v2
    assert len(seen) == len(shared)
    if mesh.startswith("ico") or mesh.startswith("ell"):
        assert nb == 0 and rwg.shape[0] * 2 == tri.shape[0] * 3  # closed: N = 1.5 T


def test_rwg_first_appearance_order():
    """Unknowns come in the order their edge was first met (CommonEdge.cpp:86-95)."""
    xyz, tri = meshgen.icosphere(2, 1.0)
    rwg, _ = O.build_rwg(tri)
    first = {}
    for t, (a, b, c) in enumerate(tri):
        for k, (u, v) in enumerate(((a, b), (b, c), (c, a))):
            first.setdefault(frozenset((u, v)), (t, k))
    keys = [first[frozenset((r[0], r[1]))] for r in rwg]
    assert keys == sorted(keys)


# ---------------------------------------------------------------- singular potentials
def _duffy_potentials(tri, r):
    """int 1/R and int (r'-r)/R over a triangle containing r in its plane,
    split at r and integrated in Duffy coordinates (1-D smooth integrals)."""
    is2 = 0.0
    is1 = np.zeros(3)
    for i in range(3):
        a, b = tri[i], tri[(i + 1) % 3]
        cr = np.linalg.norm(np.cross(a - r, b - a))
        if cr < 1e-15:
            continue
        f = lambda v: cr / np.linalg.norm((a - r) + v * (b - a))
        is2 += si.quad(f, 0.0, 1.0, epsabs=1e-14, epsrel=1e-13)[0]
        for c in range(3):
            g = lambda v, c=c: 0.5 * cr * ((a - r) + v * (b - a))[c] / np.linalg.norm((a - r) + v * (b - a))
            is1[c] += si.quad(g, 0.0, 1.0, epsabs=1e-14, epsrel=1e-13)[0]
    return is1, is2


@pytest.mark.parametrize("seed", range(6))
def test_self_potentials_vs_duffy(seed):
    rng = np.random.default_rng(seed)
    tri = rng.standard_normal((3, 3)) * 0.05
    bary, _ = O.quad_rule(7 if seed % 2 else 3)
    for b in bary:
        r = b @ tri
        is1, is2 = O.potentials(tri, r)
        d1, d2 = _duffy_potentials(tri, r)
        assert abs(is2 - d2) < 1e-10 * abs(d2)
        assert np.linalg.norm(is1 - d1) < 1e-9 * max(np.linalg.norm(d1), 1e-12)


def test_off_plane_potential_vs_2d_quadrature():
    tri = np.array([[0.0, 0.0, 0.0], [0.1, 0.0, 0.0], [0.02, 0.09, 0.0]])
    r = np.array([0.03, 0.02, 0.015])
    _, is2 = O.potentials(tri, r)
    # 2-D quadrature over the triangle with barycentric (s, t)
    e1, e2 = tri[1] - tri[0], tri[2] - tri[0]
    area2 = np.linalg.norm(np.cross(e1, e2))
    f = lambda t, s: area2 / np.linalg.norm(tri[0] + s * e1 + t * e2 - r)
    val = si.dblquad(f, 0, 1, 0, lambda s: 1 - s, epsabs=1e-12, epsrel=1e-11)[0]
    assert abs(is2 - val) < 1e-9 * val


# ---------------------------------------------------------------- fill invariants
def test_efie_reciprocity():
    xyz, tri = meshgen.icosphere(1, 0.5)
    rwg, _ = O.build_rwg(tri)
    k = O.wavenumber(3e8)
    Z = O.fill_efie(xyz, tri, rwg, k, 3)
    share = np.array([[len({r[4], r[5]} & {s[4], s[5]}) > 0 for s in rwg] for r in rwg])
    D = np.abs(Z - Z.T)
    assert D[~share].max() < 1e-13 * np.abs(Z).max()
    assert D[share].max() > 0  # the self-term rule is not symmetric, as in the reference


def test_fill_row_blocks_consistent():
    xyz, tri = meshgen.icosphere(1, 1.0)
    rwg, _ = O.build_rwg(tri)
    k = O.wavenumber(6e8)
    Z = O.fill_cfie(xyz, tri, rwg, k, 0.5, 3, 0)
    Zb = O.fill_cfie(xyz, tri, rwg, k, 0.5, 3, 0, rows=(10, 50))
    assert np.array_equal(Z[10:50], Zb)


def test_cfie_alpha_one_is_efie():
    xyz, tri = meshgen.icosphere(1, 0.5)
    rwg, _ = O.build_rwg(tri)
    k = O.wavenumber(3e8)
    Ze = O.fill_efie(xyz, tri, rwg, k, 3)
    Zc = O.fill_cfie(xyz, tri, rwg, k, 1.0, 3, 1)  # mode 1 = EFIE-style self term
    assert np.abs(Ze - Zc).max() < 1e-12 * np.abs(Ze).max()


# ---------------------------------------------------------------- linear algebra
@pytest.mark.parametrize("n", [7, 64, 150])
def test_lu_matches_lapack_getrf(n):
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    LU, piv, info = O.lu(A)
    lu_s, piv_s = sl.lu_factor(A)
    assert info == 0 and np.array_equal(piv, piv_s)
    assert np.abs(LU - lu_s).max() < 1e-12 * np.abs(lu_s).max()
    B = rng.standard_normal((n, 3)) + 1j * rng.standard_normal((n, 3))
    X = O.lu_solve(LU, piv, B)
    assert np.linalg.norm(X - sl.lu_solve((lu_s, piv_s), B)) < 1e-11 * np.linalg.norm(X)


def test_lu_zero_pivot_info():
    A = np.eye(5, dtype=complex)
    A[2, 2] = 0
    _, _, info = O.lu(A)
    assert info == 3


def test_krylov_solvers_reach_lu_solution():
    xyz, tri = meshgen.icosphere(2, 1.0)
    rwg, _ = O.build_rwg(tri)
    k = O.wavenumber(6e8)
    Z = O.fill_cfie(xyz, tri, rwg, k, 0.5, 3, 0)
    b = O.rhs(xyz, tri, rwg, 1, k, 0.5, O.inc_table([(0, 0, 0)]))[:, 0]
    x_lu = np.linalg.solve(Z, b)
    x, it, rr = O.gmres(Z, b, m=30, tol=1e-10, max_restart=50)
    assert it > 0 and rr < 1e-10
    assert np.linalg.norm(x - x_lu) < 1e-8 * np.linalg.norm(x_lu)
    x2, it2, ok = O.bicgstab(Z, b, eps=1e-8, max_iter=500)
    assert ok and np.linalg.norm(x2 - x_lu) < 1e-5 * np.linalg.norm(x_lu)


# ---------------------------------------------------------------- physics
def _sphere_chain(kind, freq, radius, subdiv, quad=3):
    xyz, tri = meshgen.icosphere(subdiv, radius)
    rwg, _ = O.build_rwg(tri)
    k = O.wavenumber(freq)
    if kind == 0:
        Z = O.fill_efie(xyz, tri, rwg, k, quad)
    else:
        Z = O.fill_cfie(xyz, tri, rwg, k, 0.5 if kind == 1 else 0.0, quad, 0)
    B = O.rhs(xyz, tri, rwg, kind, k, 0.5, O.inc_table([(0, 0, 0)]))
    X = np.linalg.solve(Z, B)
    th = np.arange(0.0, 181.0, 5.0)
    rows = []
    for ph in (0.0, 90.0):
        dirs = np.stack([np.sin(np.radians(th)) * np.cos(np.radians(ph)),
                         np.sin(np.radians(th)) * np.sin(np.radians(ph)), np.cos(np.radians(th))], 1)
        rows.append((O.rcs(xyz, tri, rwg, k, X, dirs), mie.pec_sphere_rcs_db(radius, k, th, ph)))
    return rows


@pytest.mark.parametrize("kind,tol", [(0, 0.5), (1, 0.8), (2, 1.2)])
def test_sphere_rcs_vs_mie(kind, tol):
    """PEC sphere r=0.5 m at 300 MHz, icosphere s=3 (1920 RWG): E- and H-plane cuts."""
    for rcs_o, rcs_m in _sphere_chain(kind, 3e8, 0.5, 3):
        err = np.abs(rcs_o - rcs_m)
        # compare away from the deep nulls, where dB differences blow up
        mask = rcs_m > rcs_m.max() - 15
        assert err[mask].max() < tol, (kind, err[mask].max())


def test_mie_limits():
    k = 0.05
    assert abs(mie.pec_sphere_rcs(1.0, k, 0, 0)[0] / (9 * np.pi * k ** 4) - 1) < 1e-2  # Rayleigh backscatter
    assert abs(mie.pec_sphere_rcs(1.0, 300.0, 0, 0)[0] / np.pi - 1) < 2e-2  # optical limit pi a^2
