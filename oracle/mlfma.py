"""numpy restatement of the reference MLFMA (EMSolver/MLFMA.cpp) for small
meshes, FP64 throughout.

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product package.

Steps, each following the reference routine named beside it:
  octree            MLFMA_Boundary.cpp:15-35 (cube = extent + 0.1 lambda,
                    halvings until <= threshold*lambda), MLFMA.cpp:688-875
                    (octant split by RWG edge midpoint, non-empty children
                    in (x*2+y)*2+z order, near = parent-near children closer
                    than 1.75 box lengths, far = the rest)
  samples           MLFMA.cpp:877-915,1133-1166 (L = int(kd + 6 cbrt(kd)),
                    kd = 1.732051 k a; L Gauss-Legendre theta nodes on
                    theta = pi/2 (x+1), 2L nodes on phi = pi (x+1),
                    sample k = p*L + t, weight w_phi w_theta)
  interpolation     MLFMA.cpp:917-1002 (3-point Lagrange per dimension on the
                    Gauss-Legendre abscissae, 9 weights per coarse sample)
  translation       MLFMA.cpp:1191-1222, Quadrature.h:97-149
  rad / rec         MLFMA.cpp:1244-1292,1393-1422
  coefficients      MLFMA.cpp:382-400
  matvec            MLFMA.cpp:454-686 (leaf aggregation, upward pass,
                    translation, downward pass, receive, + near matrix)
  preconditioner    MLFMA.cpp:714-741,1294-1391 (sparse approximate inverse:
                    per leaf box, the near entries between the near-box
                    unknowns within row_threshold and col_threshold leaf
                    edges of its centre, least squares by QR, one column of
                    P per own unknown; the operator becomes P Z, the
                    right-hand side P V, MLFMA.cpp:406-407,506-508)
Unknowns are kept in the original RWG order at the interface; `perm` gives
the reference's box order.
"""
import numpy as np

Z0 = 120.0 * np.pi
C0 = 299792458.0
R3 = np.array([[2 / 3, 1 / 6, 1 / 6], [1 / 6, 2 / 3, 1 / 6], [1 / 6, 1 / 6, 2 / 3]])  # Gauss3Point barycentrics


def multipoles(k, a):
    """At least 3 (the 3-point interpolation needs three nodes)."""
    kd = 1.732051 * k * a
    return max(3, int(kd + 6.0 * np.cbrt(kd)))


def lagrange(nxt, cur):
    """(loc 3 x n, weights 3 x n) of every cur abscissa on the nxt grid
    (MLFMA.cpp:957-1002: the index walk over sorted abscissae)."""
    n = len(cur)
    loc = np.zeros((3, n), dtype=np.int64)
    el = np.zeros((3, n))
    nlast = len(nxt) - 1
    nid = 2
    low, high = nxt[1], nxt[nlast - 1]
    for cid in range(n):
        xx = cur[cid]
        if xx < low:
            p = (0, 1, 2)
        elif xx > high:
            p = (nlast - 2, nlast - 1, nlast)
        else:
            nid += 1 if nxt[nid] < xx else 0
            off = 1 if (nxt[nid] - xx < xx - nxt[nid - 1]) else 0
            p = (nid + off - 2, nid + off - 1, nid + off)
        x0, x1, x2 = nxt[p[0]], nxt[p[1]], nxt[p[2]]
        loc[:, cid] = p
        el[0, cid] = (xx - x1) * (xx - x2) / ((x0 - x1) * (x0 - x2))
        el[1, cid] = (xx - x0) * (xx - x2) / ((x1 - x0) * (x1 - x2))
        el[2, cid] = (xx - x0) * (xx - x1) / ((x2 - x0) * (x2 - x1))
    return loc, el


def sph_hankel2(x, lmax):
    """h_l^(2)(x), l = 0..lmax, by the reference's upward recurrence."""
    h = np.zeros(lmax + 1, dtype=complex)
    h[0] = complex(np.sin(x), np.cos(x)) / x
    if lmax >= 1:
        h[1] = complex(np.sin(x) / x - np.cos(x), np.cos(x) / x + np.sin(x)) / x
    for n in range(2, lmax + 1):
        h[n] = ((2 * n - 1) / x) * h[n - 1] - h[n - 2]
    return h


def legendre(x, lmax):
    """P_l(x) for array x, l = 0..lmax (rows)."""
    P = np.zeros((lmax + 1,) + np.shape(x))
    P[0] = 1.0
    if lmax >= 1:
        P[1] = x
    for n in range(2, lmax + 1):
        P[n] = ((2 * n - 1) * x * P[n - 1] - (n - 1) * P[n - 2]) / n
    return P


def translation(k, D, khat, Lt):
    """T(k) = sum_{l<=min(Lt,kD)} (2l+1) (-j)^l h_l^(2)(kD) P_l(khat.Dhat)."""
    d = np.linalg.norm(D)
    kd = k * d
    lmax = int(min(Lt, kd))
    h = sph_hankel2(kd, lmax)
    P = legendre(khat @ (D / d), lmax)
    jl = np.array([1, -1j, -1, 1j])
    coef = np.array([(2 * l + 1) * jl[l % 4] * h[l] for l in range(lmax + 1)])
    return coef @ P


class Tree:
    def __init__(self, xyz, tri, rwg, freq, threshold):
        lam = C0 / freq
        self.k = 2 * np.pi * freq / C0
        mn, mx = xyz.min(0), xyz.max(0)
        center = (mn + mx) / 2
        L0 = float((mx - mn).max()) + 0.1 * lam
        delta = threshold * lam
        level, ml = 0, L0
        while ml > delta:
            ml /= 2
            level += 1
        if level < 2:
            raise ValueError("MLFMA needs at least two levels")
        self.level = level
        self.L0 = L0
        mid = (xyz[rwg[:, 0]] + xyz[rwg[:, 1]]) / 2
        # boxes per level: list of dicts(center, unknowns, parent)
        boxes = [[dict(center=center, unk=list(range(len(rwg))), parent=-1)]]
        blen = [L0]
        for l in range(1, level + 1):
            a = L0 / 2 ** l
            blen.append(a)
            cur = []
            for pi, pb in enumerate(boxes[l - 1]):
                kids = [[] for _ in range(8)]
                c = pb["center"]
                for u in pb["unk"]:
                    idx = ((0 if mid[u, 0] < c[0] else 1) * 2 + (0 if mid[u, 1] < c[1] else 1)) * 2 + \
                          (0 if mid[u, 2] < c[2] else 1)
                    kids[idx].append(u)
                for idx in range(8):
                    if kids[idx]:
                        x, y, z = idx >> 2, (idx >> 1) & 1, idx & 1
                        cc = c + a * np.array([x - 0.5, y - 0.5, z - 0.5])
                        cur.append(dict(center=cc, unk=kids[idx], parent=pi, octant=idx))
            boxes.append(cur)
        self.boxes, self.blen = boxes, blen
        # near / far (MLFMA.cpp:836-875)
        for b in boxes[1]:
            b["near"] = list(range(len(boxes[1])))
            b["far"] = []
        for l in range(2, level + 1):
            a = blen[l]
            children = [[] for _ in boxes[l - 1]]
            for i, b in enumerate(boxes[l]):
                children[b["parent"]].append(i)
            for b in boxes[l]:
                b["near"], b["far"] = [], []
                for pn in boxes[l - 1][b["parent"]]["near"]:
                    for ci in children[pn]:
                        d = np.linalg.norm(boxes[l][ci]["center"] - b["center"])
                        (b["near"] if d < 1.75 * a else b["far"]).append(ci)
        self.perm = np.array([u for b in boxes[level] for u in b["unk"]], dtype=np.int64)
        # samples per level 2..level
        self.L, self.kh, self.w, self.xt, self.xp, self.theta = {}, {}, {}, {}, {}, {}
        for l in range(2, level + 1):
            L = multipoles(self.k, blen[l])
            xt, wt = np.polynomial.legendre.leggauss(L)
            xp, wp = np.polynomial.legendre.leggauss(2 * L)
            th = np.pi / 2 * (xt + 1)
            ph = np.pi * (xp + 1)
            T, P = np.meshgrid(th, ph)  # [p, t]
            self.kh[l] = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
            self.w[l] = (wp[:, None] * wt[None, :]).reshape(-1)
            self.L[l], self.xt[l], self.xp[l], self.theta[l] = L, xt, xp, np.tile(th, 2 * L)
        # interpolation: parent level l samples from child level l+1 samples
        self.ip = {}
        for l in range(2, level):
            tloc, tel = lagrange(self.xt[l + 1], self.xt[l])
            ploc, pel = lagrange(self.xp[l + 1], self.xp[l])
            ntc = self.L[l + 1]
            K = 2 * self.L[l] ** 2
            M = np.zeros((K, 2 * self.L[l + 1] ** 2))
            row = 0
            for p in range(2 * self.L[l]):
                for t in range(self.L[l]):
                    for a in range(3):
                        for b in range(3):
                            M[row, ploc[a, p] * ntc + tloc[b, t]] += pel[a, p] * tel[b, t]
                    row += 1
            self.ip[l] = M

    def rad_rec(self, xyz, tri, rwg, alpha, subset=None):
        """rad (N x K x 3) and rec (N x K x 3) at the leaf level, original RWG
        order (only the rows in `subset` when given)."""
        l = self.level
        kh, k = self.kh[l], self.k
        n = len(rwg)
        rad = np.zeros((n, len(kh), 3), dtype=complex)
        rec = np.zeros((n, len(kh), 3), dtype=complex)
        cen = {}
        for b in self.boxes[l]:
            for u in b["unk"]:
                cen[u] = b["center"]
        for u in (range(n) if subset is None else subset):
            v1, v2, vp, vm, tp, tm = rwg[u]
            length = np.linalg.norm(xyz[v1] - xyz[v2])
            c = cen[u]
            val = np.zeros((len(kh), 3), dtype=complex)
            hr = np.zeros((len(kh), 3), dtype=complex)
            for t, vfree, s in ((tp, vp, 1.0), (tm, vm, -1.0)):
                P = xyz[tri[t]]
                nrm = np.cross(P[1] - P[0], P[2] - P[0])
                nrm /= np.linalg.norm(nrm)
                for q in range(3):
                    r = R3[q] @ P
                    rho = s * (r - xyz[vfree])
                    ph = kh @ (r - c) * k
                    val += (1 / 3) * np.exp(1j * ph)[:, None] * rho[None, :]
                    hr += (1 / 3) * np.exp(-1j * ph)[:, None] * np.cross(nrm, rho)[None, :]
            val = val - (np.sum(kh * val, 1))[:, None] * kh
            rad[u] = length / 2 * val
            rec[u] = length / 2 * (alpha * np.conj(val) + (1 - alpha) * np.cross(kh, hr))
        return rad, rec

    def near_pairs(self):
        """(rows, cols) of the near matrix, original RWG indices."""
        l = self.level
        rows, cols = [], []
        for b in self.boxes[l]:
            for nb in b["near"]:
                for f in self.boxes[l][nb]["unk"]:
                    for s in b["unk"]:
                        rows.append(f)
                        cols.append(s)
        return np.array(rows), np.array(cols)

    def far_matvec(self, x, rad, rec):
        """Far-field part of Z x (original RWG order); the patterns may carry
        any number of channels (3 vector components, a fourth scalar one)."""
        lv, k = self.level, self.k
        boxes = self.boxes
        nc = rad.shape[2]
        up = {lv: np.array([np.einsum("u,ukc->kc", x[b["unk"]], rad[b["unk"]]) for b in boxes[lv]])}
        for l in range(lv - 1, 1, -1):
            K = len(self.kh[l])
            out = np.zeros((len(boxes[l]), K, nc), dtype=complex)
            for ci, c in enumerate(boxes[l + 1]):
                p = c["parent"]
                r = boxes[l][p]["center"] - c["center"]
                ph = np.exp(-1j * k * (self.kh[l] @ r))
                out[p] += ph[:, None] * (self.ip[l] @ up[l + 1][ci])
            up[l] = out
        down = {}
        for l in range(2, lv + 1):
            K = len(self.kh[l])
            dn = np.zeros((len(boxes[l]), K, nc), dtype=complex)
            for bi, b in enumerate(boxes[l]):
                for fi in b["far"]:
                    T = translation(k, b["center"] - boxes[l][fi]["center"], self.kh[l], self.L[l])
                    dn[bi] += T[:, None] * up[l][fi]
                if l > 2:
                    p = b["parent"]
                    r = boxes[l - 1][p]["center"] - b["center"]
                    ph = np.exp(1j * k * (self.kh[l - 1] @ r))
                    A = self.ip[l - 1] * self.w[l - 1][:, None] / self.w[l][None, :]  # [kp, kc]
                    dn[bi] += A.T @ (ph[:, None] * down[l - 1][p])
            down[l] = dn
        coeff = self.w[lv] * np.sin(self.theta[lv]) * k * k * Z0 / 32
        y = np.zeros(len(x), dtype=complex)
        for bi, b in enumerate(boxes[lv]):
            u = b["unk"]
            y[u] = np.einsum("ukc,kc,k->u", rec[u], down[lv][bi], coeff)
        return y


def sai(boxes, mid, Znear, row_threshold, col_threshold, box_len):
    """Sparse approximate inverse of MLFMA::calculatePreconditioner
    (MLFMA.cpp:1294-1391) as a dense n x n matrix in RWG order.  `boxes` are
    leaf boxes (dicts with center, unk, near: indices into boxes) in box
    order, `mid` the RWG edge midpoints, Znear the near matrix (RWG order,
    zero outside the near pattern); the thresholds are in leaf edges.  For
    leaf box b: rows / cols = unknowns of its near boxes (near-list order)
    whose midpoints lie within row / col threshold of b's centre; A =
    Znear[rows][:, cols]; A = Q R (economy); column u of P (u an unknown of
    b) holds R^-1 Q[own row of u, :]^H at rows cols."""
    n = len(mid)
    P = np.zeros((n, n), dtype=complex)
    rt, ct = row_threshold * box_len, col_threshold * box_len
    for b in boxes:
        rows, cols, own = [], [], None
        for nb in b["near"]:
            nbx = boxes[nb]
            if nbx is b:
                own = len(rows)
            for u in nbx["unk"]:
                d = np.linalg.norm(mid[u] - b["center"])
                if d <= rt:
                    rows.append(u)
                if d <= ct:
                    cols.append(u)
        A = Znear[np.ix_(rows, cols)]
        Q, R = np.linalg.qr(A)
        nu = len(b["unk"])
        res = np.linalg.solve(R, Q[own:own + nu, :].conj().T)
        for j, u in enumerate(b["unk"]):
            P[cols, u] = res[:, j]
    return P
