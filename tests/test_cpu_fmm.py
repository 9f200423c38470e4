"""Pins of the single-level FMM restatement (oracle/fmm.py) that need no
device: the box grid against its defining inequalities and a brute-force
FP64 recomputation, the near/far split against the 1.75-delta distance rule,
the shift of the box centres (FMM_Boundary.cpp:43-51 leaves out the lower
bound) leaving the far product unchanged, L against hand-computed values,
and far interactions against dense CFIE entries of the CPU oracle."""
import numpy as np
import pytest

from conftest import sphere
from oracle import fmm as OF
from oracle import oracle as O
from oracle.mlfma import C0, translation


def _plate_mesh(cells, side=1.0):
    from emfa_amd import hip
    xyz, tri = hip.plate(cells, side)
    rwg, _ = hip.rwg_build(tri)
    return xyz, tri, rwg


@pytest.mark.parametrize("mesh,freq,box", [("sphere", 1.2e9, 0.4), ("sphere", 6e8, 0.25), ("plate", 1.5e9, 0.5)])
def test_grid_invariants(mesh, freq, box):
    xyz, tri, rwg = sphere(3, 0.5) if mesh == "sphere" else _plate_mesh(12)
    G = OF.Grid(xyz, tri, rwg, freq, box)
    lam = C0 / freq
    ext = (xyz.max(0) - xyz.min(0)).max()
    num = round(ext / (box * lam))
    assert abs(G.delta - ext / num) < 1e-6 * ext
    # the widened bounds hold every vertex with 0.1 delta to spare, the grid covers them
    assert np.all(G.lower <= xyz.min(0) - 0.1 * G.delta + 1e-6)
    assert np.all(np.array(G.grid) * G.delta >= (xyz.max(0) - xyz.min(0)) + 0.2 * G.delta - 1e-6)
    # every RWG midpoint lies in its box (brute force in FP64; FP32 rounding at most 1e-6)
    mid = (xyz[rwg[:, 0]] + xyz[rwg[:, 1]]) / 2
    for b in G.boxes:
        lo = G.lower + b["ix"] * G.delta
        m = mid[b["unk"]]
        assert np.all(m >= lo - 1e-6) and np.all(m <= lo + G.delta + 1e-6)
        # reference centre: (ix + 0.5) delta, the lower bound left out
        assert np.allclose(b["center"], (b["ix"] + 0.5) * G.delta, rtol=1e-6)
    # boxes in grid order, unknowns in RWG order, a permutation overall
    cells = [int((b["ix"][0] * G.grid[1] + b["ix"][1]) * G.grid[2] + b["ix"][2]) for b in G.boxes]
    assert cells == sorted(cells) and len(set(cells)) == len(cells)
    assert all(b["unk"] == sorted(b["unk"]) for b in G.boxes)
    assert sorted(G.perm.tolist()) == list(range(len(rwg)))
    # near = the 27 grid neighbours (distance < 1.75 delta), far = the rest
    for i, b in enumerate(G.boxes):
        cheb = [int(np.abs(G.boxes[j]["ix"] - b["ix"]).max()) for j in range(len(G.boxes))]
        assert b["near"] == [j for j, c in enumerate(cheb) if c <= 1]
        assert b["far"] == [j for j, c in enumerate(cheb) if c > 1]


def test_multipoles_follow_the_reference_formula():
    # kd = 1.732051 k delta, L = int(kd + 6 cbrt(kd)): hand-computed cases
    xyz, tri, rwg = sphere(2, 0.5)
    for freq, box in ((3e8, 0.4), (1.2e9, 0.4), (6e8, 0.25)):
        G = OF.Grid(xyz, tri, rwg, freq, box)
        kd = 1.732051 * (2 * np.pi * freq / C0) * G.delta
        assert G.L == int(kd + 6 * np.cbrt(kd))
    assert len(G.kh) == 2 * G.L * G.L
    assert abs(G.w.sum() - 4.0) < 1e-12  # Gauss-Legendre weights on [-1, 1] x [-1, 1]


def test_common_centre_shift_leaves_the_far_product_unchanged():
    """FMM_Boundary.cpp:43-51 places box centres without the lower bound.
    Expanding about the true centres (lower + (ix + 0.5) delta) instead gives
    the same far matvec: the shift's phase cancels between the radiation and
    receive patterns while the translations depend on centre differences."""
    xyz, tri, rwg = sphere(2, 0.5)
    G = OF.Grid(xyz, tri, rwg, 1.2e9, 0.4)
    n = len(rwg)
    rng = np.random.default_rng(1)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    rad, rec = G.rad_rec(xyz, tri, rwg, 0.5)
    y_ref = G.far_matvec(x, rad, rec)
    for b in G.boxes:
        b["center"] = G.lower + (b["ix"] + 0.5) * G.delta
    rad2, rec2 = G.rad_rec(xyz, tri, rwg, 0.5)
    y_true = G.far_matvec(x, rad2, rec2)
    assert np.abs(rad2 - rad).max() > 1e-3  # the patterns themselves differ
    # equal up to the FP32 rounding of the reference's centres (their
    # differences enter the translations)
    assert np.linalg.norm(y_true - y_ref) / np.linalg.norm(y_ref) < 1e-6


def test_far_interactions_match_dense_entries():
    """rec . T . rad of single RWG pairs against dense CFIE entries, with
    twice the reference's samples per angle: the reference's own L x 2L
    Gauss-Legendre samples do not integrate the translation's top orders
    (errors up to order one on small far entries at k D > L; the operator
    test below measures what that leaves in a matvec).  The formulas are
    the reference's; a dropped conjugate fails the check."""
    xyz, tri, rwg = sphere(3, 0.5)
    f, alpha = 1.2e9, 0.5
    G = OF.Grid(xyz, tri, rwg, f, 0.4, oversample=2)
    k = G.k
    boxes = G.boxes
    rng = np.random.default_rng(2)
    pairs = []
    while len(pairs) < 6:
        bi = int(rng.integers(len(boxes)))
        far = [fi for fi in boxes[bi]["far"]
               if np.linalg.norm(boxes[bi]["center"] - boxes[fi]["center"]) >= 3.9 * G.delta]
        if far:
            pairs.append((bi, far[int(rng.integers(len(far)))]))
    us = sorted({boxes[b]["unk"][0] for p in pairs for b in p})
    rad, rec = G.rad_rec(xyz, tri, rwg, alpha, subset=us)
    errs = []
    for bi, fi in pairs:
        m, nn = boxes[bi]["unk"][0], boxes[fi]["unk"][0]
        Tk = translation(k, boxes[bi]["center"] - boxes[fi]["center"], G.kh, G.L)
        zm = np.sum(G.coeff[:, None] * rec[m] * Tk[:, None] * rad[nn])
        zd = O.fill_cfie(xyz, tri, rwg, k, alpha, 3, 0, rows=(m, m + 1))[0, nn]
        errs.append(abs(zm - zd) / abs(zd))
    assert max(errs) < 3e-2 and np.median(errs) < 1e-2
    # a receive pattern without the conjugate (a plausible slip) misses the entries
    Tk = translation(k, boxes[pairs[0][0]]["center"] - boxes[pairs[0][1]]["center"], G.kh, G.L)
    m, nn = boxes[pairs[0][0]]["unk"][0], boxes[pairs[0][1]]["unk"][0]
    bad = np.sum(G.coeff[:, None] * np.conj(rec[m]) * Tk[:, None] * rad[nn])
    zd = O.fill_cfie(xyz, tri, rwg, k, alpha, 3, 0, rows=(m, m + 1))[0, nn]
    assert abs(bad - zd) / abs(zd) > 0.3


def test_operator_close_to_dense():
    """Near entries plus the restated far part (the reference's sampling)
    against the dense CFIE matrix: 3.2e-3 at 600 MHz (6 x 6 x 6 boxes); at
    1.2 GHz on the same mesh (11 x 11 x 11 boxes, more pairs at k D > L)
    the reference's algorithm leaves 3.4e-2."""
    xyz, tri, rwg = sphere(3, 0.5)
    G = OF.Grid(xyz, tri, rwg, 6e8, 0.4)
    n = len(rwg)
    Z = O.fill_cfie(xyz, tri, rwg, G.k, 0.5, 3, 0)
    rows, cols = G.near_pairs()
    mask = np.zeros((n, n), dtype=bool)
    mask[rows, cols] = True
    rng = np.random.default_rng(3)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    rad, rec = G.rad_rec(xyz, tri, rwg, 0.5)
    y = (Z * mask) @ x + G.far_matvec(x, rad, rec)
    yd = Z @ x
    assert np.linalg.norm(y - yd) / np.linalg.norm(yd) < 1e-2
