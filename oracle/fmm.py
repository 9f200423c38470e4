"""numpy restatement of the reference single-level FMM (EMSolver/FMM.cpp,
EMComponent/FMM_Boundary.cpp) for small meshes.

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product package.

Steps, each following the reference routine named beside it:
  box grid          FMM_Boundary.cpp:15-33, in FP32 as the reference computes
                    it: delta = box length * lambda, num = int(extent / delta
                    + 0.5), delta = extent / num, bounds widened by 0.1 delta,
                    boxes per axis int((size + 0.9 delta) / delta)
  box of an RWG     FMM_Boundary.cpp:35-41 on the edge midpoint (v1 + v2) / 2,
                    index (x * ny + y) * nz + z
  box centre        FMM_Boundary.cpp:43-51: ((x + 0.5) delta, ...) -- without
                    the lower bound, so every centre is shifted by the same
                    vector (harmless: the phase cancels between radiation and
                    receive patterns; kept so the patterns are the reference's)
  boxes, near/far   FMM.cpp:157-206: non-empty boxes in grid order, unknowns
                    in RWG order, near = centres closer than 1.75 delta, far =
                    every other box
  samples           FMM.cpp:47-50,322-347: L = int(kd + 6 cbrt(kd)),
                    kd = 1.732051 k delta (FP32); L Gauss-Legendre theta
                    nodes, 2L phi nodes, sample k = p * L + t
  translation       FMM.cpp:349-381 (oracle/mlfma.py translation)
  rad / rec         FMM.cpp:383-466 (same patterns as the MLFMA leaf level)
  coefficients      FMM.cpp:208-220
  matvec            FMM.cpp:270-320 (aggregation, translation from every far
                    box, receive, + near matrix)
FP64 everywhere except the grid and L, which follow the reference's FP32.
Unknowns stay in the original RWG order at the interface; `perm` gives the
reference's box order.
"""
import numpy as np

from oracle.mlfma import C0, R3, Z0, translation

f32 = np.float32


class Grid:
    def __init__(self, xyz, tri, rwg, freq, box_len, oversample=1):
        """oversample > 1 multiplies the theta and phi sample counts (the
        translation stays truncated at L): a check of the formulas, not the
        reference's sampling."""
        self.k = 2 * np.pi * freq / C0
        v = xyz.astype(f32)
        mn, mx = v.min(0), v.max(0)
        lam = f32(C0 / freq)
        delta = f32(box_len) * lam
        ext = mx - mn
        max_len = max(ext[0], ext[1], ext[2])
        num = int(f32(max_len / delta) + f32(0.5))
        if num < 1:
            raise ValueError("box length above twice the mesh extent")
        delta = f32(max_len / f32(num))
        off = f32(f32(0.1) * delta)
        mn = (mn - off).astype(f32)
        mx = (mx + off).astype(f32)
        size = (mx - mn).astype(f32)
        nb = [int(f32(f32(size[c] + f32(f32(0.9) * delta)) / delta)) for c in range(3)]
        mid = ((v[rwg[:, 0]] + v[rwg[:, 1]]) / f32(2)).astype(f32)
        ix = ((mid - mn) / delta).astype(f32).astype(np.int64)
        assert np.all(ix >= 0) and np.all(ix < np.array(nb))
        cell = (ix[:, 0] * nb[1] + ix[:, 1]) * nb[2] + ix[:, 2]
        self.delta, self.grid, self.lower = float(delta), nb, mn.astype(float)
        self.cell_of = cell
        self.boxes = []
        for c in np.unique(cell):
            x, y, z = int(c // (nb[1] * nb[2])), int((c // nb[2]) % nb[1]), int(c % nb[2])
            cen = np.array([f32(f32(x + 0.5) * delta), f32(f32(y + 0.5) * delta), f32(f32(z + 0.5) * delta)],
                           dtype=float)
            self.boxes.append(dict(ix=np.array([x, y, z]), center=cen, unk=np.nonzero(cell == c)[0].tolist()))
        cen = np.array([b["center"] for b in self.boxes], dtype=f32)
        lim = f32(f32(1.75) * delta)
        for i, b in enumerate(self.boxes):
            d = np.sqrt(np.sum((cen - cen[i]) ** 2, 1, dtype=f32), dtype=f32)
            b["near"] = np.nonzero(d < lim)[0].tolist()
            b["far"] = np.nonzero(~(d < lim))[0].tolist()
        self.perm = np.array([u for b in self.boxes for u in b["unk"]], dtype=np.int64)
        kd = f32(f32(1.732051) * f32(self.k)) * delta
        self.L = int(f32(kd + f32(6) * np.cbrt(kd, dtype=f32)))
        L = self.L
        nt = L * oversample
        xt, wt = np.polynomial.legendre.leggauss(nt)
        xp, wp = np.polynomial.legendre.leggauss(2 * nt)
        th, ph = np.pi / 2 * (xt + 1), np.pi * (xp + 1)
        T, P = np.meshgrid(th, ph)  # [p, t]
        self.kh = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
        self.w = (wp[:, None] * wt[None, :]).reshape(-1)
        self.theta = np.tile(th, 2 * nt)
        self.coeff = self.w * np.sin(self.theta) * self.k ** 2 * Z0 / 32

    def center_of(self):
        cen = np.zeros((len(self.cell_of), 3))
        for b in self.boxes:
            cen[b["unk"]] = b["center"]
        return cen

    def rad_rec(self, xyz, tri, rwg, alpha, subset=None):
        """rad (N x K x 3) and rec (N x K x 3), original RWG order (only the
        rows in `subset` when given)."""
        kh, k = self.kh, self.k
        n = len(rwg)
        rad = np.zeros((n, len(kh), 3), dtype=complex)
        rec = np.zeros((n, len(kh), 3), dtype=complex)
        cen = self.center_of()
        for u in (range(n) if subset is None else subset):
            v1, v2, vp, vm, tp, tm = rwg[u]
            length = np.linalg.norm(xyz[v1] - xyz[v2])
            val = np.zeros((len(kh), 3), dtype=complex)
            hr = np.zeros((len(kh), 3), dtype=complex)
            for t, vfree, s in ((tp, vp, 1.0), (tm, vm, -1.0)):
                P = xyz[tri[t]]
                nrm = np.cross(P[1] - P[0], P[2] - P[0])
                nrm /= np.linalg.norm(nrm)
                for q in range(3):
                    r = R3[q] @ P
                    rho = s * (r - xyz[vfree])
                    ph = kh @ (r - cen[u]) * k
                    val += (1 / 3) * np.exp(1j * ph)[:, None] * rho[None, :]
                    hr += (1 / 3) * np.exp(-1j * ph)[:, None] * np.cross(nrm, rho)[None, :]
            val = val - (np.sum(kh * val, 1))[:, None] * kh
            rad[u] = length / 2 * val
            rec[u] = length / 2 * (alpha * np.conj(val) + (1 - alpha) * np.cross(kh, hr))
        return rad, rec

    def near_pairs(self):
        """(rows, cols) of the near matrix, original RWG indices."""
        rows, cols = [], []
        for b in self.boxes:
            for nb in b["near"]:
                for f in self.boxes[nb]["unk"]:
                    for s in b["unk"]:
                        rows.append(f)
                        cols.append(s)
        return np.array(rows), np.array(cols)

    def far_matvec(self, x, rad, rec):
        """Far part of Z x (original RWG order), FMM.cpp:270-320."""
        sk = np.array([np.einsum("u,ukc->kc", x[b["unk"]], rad[b["unk"]]) for b in self.boxes])
        y = np.zeros(len(x), dtype=complex)
        for bi, b in enumerate(self.boxes):
            gk = np.zeros((len(self.kh), 3), dtype=complex)
            for fi in b["far"]:
                T = translation(self.k, b["center"] - self.boxes[fi]["center"], self.kh, self.L)
                gk += T[:, None] * sk[fi]
            u = b["unk"]
            y[u] = np.einsum("ukc,kc,k->u", rec[u], gk, self.coeff)
        return y
