"""Pins of the MLFMA restatement (oracle/mlfma.py) that need no device:
the plane-wave expansion of the Green's function through the translation
operator and the level's samples (addition theorem), exactness of the
3-point Lagrange interpolation on quadratics, the octree invariants, and
far interactions of well-separated boxes against dense CFIE entries of the
CPU oracle."""
import numpy as np
import pytest

from conftest import sphere
from oracle import mlfma as OM
from oracle import oracle as O


def test_addition_theorem_through_translation_and_samples():
    # exp(-jkR)/R = (-jk/4pi) (pi^2/2) sum_q W_q sin(theta_q) exp(-jk khat.d) T_L(khat, D)
    k = 2 * np.pi  # lambda = 1
    rng = np.random.default_rng(0)
    a = 0.5
    L = OM.multipoles(k, a)
    xt, wt = np.polynomial.legendre.leggauss(L)
    xp, wp = np.polynomial.legendre.leggauss(2 * L)
    th, ph = np.pi / 2 * (xt + 1), np.pi * (xp + 1)
    T, P = np.meshgrid(th, ph)
    kh = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
    W = (wp[:, None] * wt[None, :]).reshape(-1) * np.sin(T).reshape(-1) * np.pi ** 2 / 2
    errs = []
    for D in ([2 * a, 0, 0], [2 * a, 2 * a, a], [3 * a, -a, 2 * a], [0, 0, -2 * a]):
        D = np.array(D, dtype=float)
        Tk = OM.translation(k, D, kh, L)
        for _ in range(5):
            rm = rng.uniform(-a / 2, a / 2, 3)  # receiver about the receiving centre
            rn = rng.uniform(-a / 2, a / 2, 3)  # source about the source centre
            R = np.linalg.norm(D + rm - rn)
            exact = np.exp(-1j * k * R) / R
            approx = (-1j * k / (4 * np.pi)) * np.sum(W * np.exp(-1j * k * (kh @ (rm - rn))) * Tk)
            errs.append(abs(approx - exact) / abs(exact))
    # truncation error of L = kd + 6 cbrt(kd): worst at box corners two edges apart
    assert np.max(errs) < 5e-2 and np.median(errs) < 2e-3
    # a wrong Hankel kind (h^(1)) or sign of (-j)^l fails the identity
    Dm = np.array([2 * a, 0, 0])
    Tbad = np.conj(OM.translation(k, Dm, kh, L))
    R = np.linalg.norm(Dm)
    bad = (-1j * k / (4 * np.pi)) * np.sum(W * Tbad)
    assert abs(bad - np.exp(-1j * k * R) / R) / abs(np.exp(-1j * k * R) / R) > 0.5


def test_lagrange_exact_on_quadratics():
    nxt = np.polynomial.legendre.leggauss(9)[0]
    cur = np.polynomial.legendre.leggauss(14)[0]
    loc, el = OM.lagrange(nxt, cur)
    f = lambda x: 3 * x * x - 2 * x + 0.5
    approx = np.sum(el * f(nxt[loc]), 0)
    assert np.abs(approx - f(cur)).max() < 1e-12
    assert np.abs(el.sum(0) - 1).max() < 1e-12
    # stencils are three consecutive nodes around each abscissa
    assert np.all(np.diff(loc, axis=0) == 1)


def test_octree_invariants():
    xyz, tri, rwg = sphere(3, 0.5)
    T = OM.Tree(xyz, tri, rwg, 1.2e9, 0.25)
    n = len(rwg)
    assert sorted(T.perm.tolist()) == list(range(n))
    lam = OM.C0 / 1.2e9
    assert T.blen[T.level] <= 0.25 * lam < T.blen[T.level - 1]
    mid = (xyz[rwg[:, 0]] + xyz[rwg[:, 1]]) / 2
    for l in range(2, T.level + 1):
        a = T.blen[l]
        for b in T.boxes[l]:
            assert np.all(np.abs(mid[b["unk"]] - b["center"]) <= a / 2 + 1e-12)
            for nb in b["near"]:
                assert np.abs(T.boxes[l][nb]["center"] - b["center"]).max() <= a * 1.0001
            for fb in b["far"]:
                assert np.abs(T.boxes[l][fb]["center"] - b["center"]).max() >= 2 * a * 0.9999


def test_far_interactions_of_separated_boxes_match_dense_entries():
    xyz, tri, rwg = sphere(4, 0.5)
    f, alpha = 8e8, 0.5
    T = OM.Tree(xyz, tri, rwg, f, 0.25)
    lv, k = T.level, T.k
    boxes = T.boxes[lv]
    rng = np.random.default_rng(2)
    pairs = []
    while len(pairs) < 6:
        bi = int(rng.integers(len(boxes)))
        far = [fi for fi in boxes[bi]["far"]
               if np.linalg.norm(boxes[bi]["center"] - boxes[fi]["center"]) >= 3.9 * T.blen[lv]]
        if far:
            pairs.append((bi, far[int(rng.integers(len(far)))]))
    us = sorted({boxes[b]["unk"][0] for p in pairs for b in p})
    rad, rec = T.rad_rec(xyz, tri, rwg, alpha, subset=us)
    coeff = T.w[lv] * np.sin(T.theta[lv]) * k * k * OM.Z0 / 32
    errs = []
    for bi, fi in pairs:
        m, nn = boxes[bi]["unk"][0], boxes[fi]["unk"][0]
        Tk = OM.translation(k, boxes[bi]["center"] - boxes[fi]["center"], T.kh[lv], T.L[lv])
        zm = np.sum(coeff[:, None] * rec[m] * Tk[:, None] * rad[nn])
        zd = O.fill_cfie(xyz, tri, rwg, k, alpha, 3, 0, rows=(m, m + 1))[0, nn]
        errs.append(abs(zm - zd) / abs(zd))
    # RWG supports reach past their leaf boxes, so a few per cent at worst
    assert max(errs) < 3e-2 and np.median(errs) < 1e-2


def test_sparse_approximate_inverse_columns_are_least_squares_solutions():
    """MLFMA::calculatePreconditioner (MLFMA.cpp:1294-1391) restated with QR:
    each column of P equals the least-squares solution (SVD-based lstsq) of
    A m = e_u over the box's rows and columns, and the row / column sets
    follow the thresholds."""
    xyz, tri, rwg = sphere(2, 0.5)
    f, alpha = 1.2e9, 0.5
    T = OM.Tree(xyz, tri, rwg, f, 0.25)
    lv = T.level
    boxes = T.boxes[lv]
    n = len(rwg)
    Z = O.fill_cfie(xyz, tri, rwg, T.k, alpha, 3, 0)
    rows, cols = T.near_pairs()
    Zn = np.zeros_like(Z)
    Zn[rows, cols] = Z[rows, cols]
    mid = (xyz[rwg[:, 0]] + xyz[rwg[:, 1]]) / 2
    a = T.blen[lv]
    P = OM.sai(boxes, mid, Zn, 0.9, 0.87, a)
    checked = 0
    for b in boxes[::3]:
        near_u = [u for nb in b["near"] for u in boxes[nb]["unk"]]
        d = np.linalg.norm(mid[near_u] - b["center"], axis=1)
        r = [u for u, x in zip(near_u, d) if x <= 0.9 * a]
        c = [u for u, x in zip(near_u, d) if x <= 0.87 * a]
        A = Zn[np.ix_(r, c)]
        for u in b["unk"]:
            e = np.zeros(len(r), dtype=complex)
            e[r.index(u)] = 1
            m = np.linalg.lstsq(A, e, rcond=None)[0]
            assert np.abs(P[c, u] - m).max() <= 1e-8 * max(1.0, np.abs(m).max())
            # nothing outside the column set
            outside = np.setdiff1d(np.arange(n), c)
            assert np.all(P[outside, u] == 0)
            checked += 1
    assert checked > 20
    # a transposed instead of conjugated Q (a plausible slip) is not the least-squares solution
    b = boxes[0]
    near_u = [u for nb in b["near"] for u in boxes[nb]["unk"]]
    d = np.linalg.norm(mid[near_u] - b["center"], axis=1)
    r = [u for u, x in zip(near_u, d) if x <= 0.9 * a]
    c = [u for u, x in zip(near_u, d) if x <= 0.87 * a]
    Q, R = np.linalg.qr(Zn[np.ix_(r, c)])
    own = r.index(b["unk"][0])
    bad = np.linalg.solve(R, Q[own, :].T)
    assert np.abs(bad - P[c, b["unk"][0]]).max() > 1e-3 * np.abs(P[c, b["unk"][0]]).max()
