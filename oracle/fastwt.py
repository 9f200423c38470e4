"""numpy restatement of the reference IP-SIE fast solver (EMSolver/FastWT.cpp,
the half-RWG interior-penalty surface integral equation on the MLFMA) for
small meshes, FP64 throughout.

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product package.

Inputs are the half-RWG list of ContourEdge (hrwg n x 6: v1, v2, vx,
triangle, edge number, type 0 conformal / 1 nonconformal; FP32 lengths; the
contour lines of the nonconformal edges), in the reference's order, which
the product builder reproduces bit for bit (tests/test_cpu_fastwt.py).

Steps, each following the reference routine named beside it:
  near entry      FastWT.cpp:734-784 (fillNearZbyColumn), for field f and
                  source s:
                    z   = l_s l_f K(f, s), K the CFIE triangle-pair kernel
                          with both signs +1 (cfieZSingular on one triangle),
                          through the C oracle's orc_fill_hrwg;
                    ip  = stab(l_f) for f = s, and for two conformal halves
                          of one edge on different triangles; stab(line) for
                          two nonconformal edges when s's triangle owns one of
                          f's contour lines (its first such line); else 0;
                          stab(len) = 2 len beta Z0 / k (FastWT.cpp:1526-1530)
                    line = -(l_f C(T_f; s.v1, s.v2) + l_s C(T_s; f.v1, f.v2)),
                          C(T; a, b) = Z0 / (4 pi j k) |b - a|
                          sum_p sum_q glw3[p] w3[q] e^{-jkr}/r,
                          r = |a + gl3[p] (b - a) - g_q(T)| (FastWT.cpp:1494-1524)
                    entry = z + alpha (line + ip)
  beta            FastWT.cpp:48: stabilization_factor |log10(average_size)|
  tree, samples   as the MLFMA (FastWT.cpp:167-400 repeats MLFMA.cpp's text),
                  unknowns binned by the half-RWG edge midpoint (v1 + v2) / 2
  patterns        FastWT.cpp:583-653 per half-RWG about its leaf centre:
                    rad  = l/2 [sum_p w3 (g_p - vx) e^{+jk khat.(g_p - c)}]_transverse
                    rec  = l/2 (alpha conj(rad') + (1 - alpha) khat x
                           sum_p w3 (n x (g_p - vx)) e^{-jk khat.(g_p - c)})
                    prad = l sum_p glw3[p] e^{+jk khat.(gl_p - c)} on the edge
                    prec = conj(prad)
  coefficients    FastWT.cpp:972-990: coeff = w sin(theta) k^2 Z0 / 32,
                  pcoeff = w sin(theta) alpha Z0 / 32
  matvec          FastWT.cpp:1006-1267: the vector patterns and the scalar
                  line pattern take the same aggregation, interpolation,
                  translation and anterpolation; far(u) = sum_k coeff rec.d
                  + sum_k pcoeff prec pd; y = near x + far
  excitation      FastWT.cpp:786-806,1475-1492: V_u = l_u / 2 sum_a w3
                  (alpha rho_a.E + (1 - alpha) rho_a.(n x H)) e^{-jk g_a.k_inc}
  RCS             FastWT.cpp:1393-1420: Es = sum_u I_u l_u sum_g w3 (g - vx)
                  e^{+jk g.s}; sigma = Z0^2 k^2 |s x (s x Es) / 2|^2 / 4 pi (dB)
  surface current FastWT.cpp:1426-1472 (per triangle, its half-RWGs in list order)
Unknowns stay in half-RWG order at the interface; the tree's `perm` is the
reference's box order.
"""
import numpy as np

from oracle import oracle as O
from oracle.mlfma import C0, R3, Z0, Tree

GL3 = np.array([0.1127016654, 0.5, 0.8872983346])  # Quadrature.h:17-18, [0, 1]
GLW3 = np.array([0.2777777777775, 0.444444444444445, 0.2777777777775])


def beta(stabilization_factor, average_size):
    return stabilization_factor * abs(np.log10(average_size))


def stabilization(length, beta_, k):
    return 2.0 * length * beta_ * Z0 / k


def line_integral(T, a, b, k):
    """|b - a| sum_p sum_q glw3[p] w3[q] e^{-jkr} / r over the three edge
    points a + gl3[p] (b - a) and the triangle's three Gauss points."""
    g = R3 @ T
    pts = a[None, :] + GL3[:, None] * (b - a)[None, :]
    r = np.linalg.norm(pts[:, None, :] - g[None, :, :], axis=2)
    return np.linalg.norm(a - b) * np.sum(GLW3[:, None] * (1.0 / 3.0) * np.exp(-1j * k * r) / r)


def consistency(T, a, b, k):
    return Z0 / (k * 4.0 * np.pi * 1j) * line_integral(T, a, b, k)


class Contour:
    """Half-RWG list with its contour lines (the product builder's or a
    golden fixture's arrays)."""

    def __init__(self, hrwg, length, line_ptr, line_tri, line_seg):
        self.hrwg = np.asarray(hrwg, dtype=np.int64)
        self.length = np.asarray(length, dtype=np.float64)
        self.line_ptr = np.asarray(line_ptr, dtype=np.int64)
        self.line_tri = np.asarray(line_tri, dtype=np.int64)
        self.line_seg = np.asarray(line_seg, dtype=np.float64).reshape(-1, 6)

    @property
    def n(self):
        return len(self.hrwg)

    def line_length(self, f, tri):
        """ContourEdge::getContourLineLength(f's edge number, tri): the FP32
        length of the first contour line of nonconformal f on triangle tri."""
        e = self.hrwg[f, 4]
        for j in range(self.line_ptr[e], self.line_ptr[e + 1]):
            if self.line_tri[j] == tri:
                s = self.line_seg[j].astype(np.float32)
                d = s[3:] - s[:3]
                return float(np.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]))
        return None


def ip_term(C, f, s, beta_, k):
    hf, hs = C.hrwg[f], C.hrwg[s]
    if hf[3] == hs[3]:
        return stabilization(C.length[f], beta_, k) if f == s else 0.0
    if hf[5] == 0 and hs[5] == 0 and hf[4] == hs[4]:
        return stabilization(C.length[f], beta_, k)
    if hf[5] == 1 and hs[5] == 1:
        ln = C.line_length(f, hs[3])
        if ln is not None:
            return stabilization(ln, beta_, k)
    return 0.0


def line_term(xyz, tri, C, f, s, k):
    hf, hs = C.hrwg[f], C.hrwg[s]
    Tf, Ts = xyz[tri[hf[3]]], xyz[tri[hs[3]]]
    return -(C.length[f] * consistency(Tf, xyz[hs[0]], xyz[hs[1]], k) +
             C.length[s] * consistency(Ts, xyz[hf[0]], xyz[hf[1]], k))


def entries(xyz, tri, C, k, alpha, beta_, pairs):
    """Near-entry formula of FastWT.cpp:734-784 for the (field, source)
    pairs given, one pair at a time."""
    out = []
    for f, s in pairs:
        zs = O.fill_hrwg(xyz, tri, C.hrwg, C.length, k, alpha, 3, 0, rows=(f, f + 1))[0, s]
        out.append(zs + alpha * (line_term(xyz, tri, C, f, s, k) + ip_term(C, f, s, beta_, k)))
    return np.array(out)


def _ip_matrix(C, k, beta_):
    """Interior-penalty part for every pair, as a dict {(f, s): value}."""
    stab = lambda ln: stabilization(ln, beta_, k)
    L = C.length
    ip = {(u, u): stab(L[u]) for u in range(C.n)}
    by_edge = {}
    for u in np.where(C.hrwg[:, 5] == 0)[0]:
        by_edge.setdefault(int(C.hrwg[u, 4]), []).append(int(u))
    for us in by_edge.values():
        for f in us:
            for s in us:
                if C.hrwg[f, 3] != C.hrwg[s, 3]:
                    ip[(f, s)] = stab(L[f])
    nc = [int(u) for u in np.where(C.hrwg[:, 5] == 1)[0]]
    for f in nc:
        for s in nc:
            if C.hrwg[f, 3] != C.hrwg[s, 3]:
                ln = C.line_length(f, C.hrwg[s, 3])
                if ln is not None:
                    ip[(f, s)] = stab(ln)
    return ip


def _gauss_and_edge_points(xyz, tri, C):
    g = np.einsum("qa,uac->uqc", R3, xyz[tri[C.hrwg[:, 3]]])          # n x 3 Gauss points
    a, b = xyz[C.hrwg[:, 0]], xyz[C.hrwg[:, 1]]
    e = a[:, None, :] + GL3[None, :, None] * (b - a)[:, None, :]       # n x 3 edge points
    return g, e, np.linalg.norm(a - b, axis=1)


def _pot(g, e, elen, t, s, k):
    """|edge s| sum_p sum_q glw3[p] w3[q] G(e_s[p], g_t[q]) for index arrays t, s."""
    r = np.linalg.norm(e[s][:, :, None, :] - g[t][:, None, :, :], axis=3)   # pair, p, q
    return elen[s] * np.einsum("p,ipq->i", GLW3 / 3.0, np.exp(-1j * k * r) / r)


def parts_at(xyz, tri, C, k, beta_, rows, cols, alpha=1.0, mode=0):
    """(Zs, line, ip) of the near-entry formula at the pairs (rows[i], cols[i]),
    vectorised in chunks."""
    rows, cols = np.asarray(rows, dtype=np.int64), np.asarray(cols, dtype=np.int64)
    Zd = O.fill_hrwg(xyz, tri, C.hrwg, C.length, k, alpha, 3, mode)
    zs = Zd[rows, cols]
    del Zd
    g, e, elen = _gauss_and_edge_points(xyz, tri, C)
    cst = Z0 / (k * 4.0 * np.pi * 1j)
    L = C.length
    line = np.zeros(len(rows), dtype=complex)
    for i0 in range(0, len(rows), 200000):
        f, s = rows[i0:i0 + 200000], cols[i0:i0 + 200000]
        line[i0:i0 + 200000] = -cst * (L[f] * _pot(g, e, elen, f, s, k) + L[s] * _pot(g, e, elen, s, f, k))
    ipd = _ip_matrix(C, k, beta_)
    ip = np.array([ipd.get((int(f), int(s)), 0.0) for f, s in zip(rows, cols)])
    return zs, line, ip


def dense_parts(xyz, tri, C, k, beta_, alpha=1.0, mode=0):
    """The three parts of the near-entry formula for every pair (n x n each):
    (Zs, line, ip) with entry = Zs + alpha (line + ip); Zs is the surface
    part at the given alpha."""
    n = C.n
    f, s = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    zs, line, ip = parts_at(xyz, tri, C, k, beta_, f.ravel(), s.ravel(), alpha, mode)
    return zs.reshape(n, n), line.reshape(n, n), ip.reshape(n, n)


def near_entries(xyz, tri, C, T, k, alpha, beta_, mode=0):
    """The operator's near matrix (dense n x n, zero off the near pattern)."""
    rows, cols = T.near_pairs()
    zs, line, ip = parts_at(xyz, tri, C, k, beta_, rows, cols, alpha, mode)
    Z = np.zeros((C.n, C.n), dtype=complex)
    Z[rows, cols] = zs + alpha * (line + ip)
    return Z


def dense(xyz, tri, C, k, alpha, beta_, mode=0):
    """Every pair's near-entry formula (n x n)."""
    Zs, line, ip = dense_parts(xyz, tri, C, k, beta_, alpha, mode)
    return Zs + alpha * (line + ip)


def tree(xyz, tri, C, freq, threshold):
    """The MLFMA octree over the half-RWG edge midpoints."""
    return Tree(xyz, tri, C.hrwg, freq, threshold)


def patterns(T, xyz, tri, C, alpha):
    """rad, rec (n x K x 4): channels 0-2 the vector patterns, channel 3 the
    line pattern prad / prec scaled by pcoeff / coeff = alpha / k^2 so that
    one coefficient serves all four channels."""
    l = T.level
    kh, k = T.kh[l], T.k
    K = len(kh)
    cen = {}
    for b in T.boxes[l]:
        for u in b["unk"]:
            cen[u] = b["center"]
    rad = np.zeros((C.n, K, 4), dtype=complex)
    rec = np.zeros((C.n, K, 4), dtype=complex)
    for u in range(C.n):
        v1, v2, vx, t = C.hrwg[u, :4]
        P = xyz[tri[t]]
        nrm = np.cross(P[1] - P[0], P[2] - P[0])
        nrm /= np.linalg.norm(nrm)
        c = cen[u]
        val = np.zeros((K, 3), dtype=complex)
        hr = np.zeros((K, 3), dtype=complex)
        for q in range(3):
            g = R3[q] @ P
            rho = g - xyz[vx]
            ph = k * (kh @ (g - c))
            val += (1 / 3) * np.exp(1j * ph)[:, None] * rho[None, :]
            hr += (1 / 3) * np.exp(-1j * ph)[:, None] * np.cross(nrm, rho)[None, :]
        val = val - np.sum(kh * val, 1)[:, None] * kh
        hl = C.length[u] / 2
        rad[u, :, :3] = hl * val
        rec[u, :, :3] = hl * (alpha * np.conj(val) + (1 - alpha) * np.cross(kh, hr))
        lr = np.zeros(K, dtype=complex)
        for p in range(3):
            gl = xyz[v1] + GL3[p] * (xyz[v2] - xyz[v1])
            lr += GLW3[p] * np.exp(1j * k * (kh @ (gl - c)))
        rad[u, :, 3] = C.length[u] * lr
        rec[u, :, 3] = (alpha / (k * k)) * C.length[u] * np.conj(lr)
    return rad, rec


def near_mask(T, n):
    rows, cols = T.near_pairs()
    m = np.zeros((n, n), dtype=bool)
    m[rows, cols] = True
    return m


def matvec(T, Znear, rad, rec, x):
    """y = near x + far x in half-RWG order (Znear dense, zero off the near pattern)."""
    return Znear @ x + T.far_matvec(x, rad, rec)


def rhs(xyz, tri, C, k, alpha, inc):
    """n x nrhs excitation (inc: nrhs x 9 = k, e, h per incidence)."""
    inc = np.atleast_2d(inc)
    B = np.zeros((C.n, len(inc)), dtype=complex)
    for u in range(C.n):
        vx, t = C.hrwg[u, 2], C.hrwg[u, 3]
        P = xyz[tri[t]]
        nrm = np.cross(P[1] - P[0], P[2] - P[0])
        nrm /= np.linalg.norm(nrm)
        for j, c in enumerate(inc):
            kv, ev, hv = c[:3], c[3:6], c[6:9]
            ve = vm = 0j
            for a in range(3):
                g = R3[a] @ P
                rho = g - xyz[vx]
                G = np.exp(-1j * k * (g @ kv))
                ve += (1 / 3) * (rho @ ev) * G
                vm += (1 / 3) * (rho @ np.cross(nrm, hv)) * G
            B[u, j] = 0.5 * C.length[u] * (alpha * ve + (1 - alpha) * vm)
    return B


def rcs(xyz, tri, C, k, I, dirs):
    """Bistatic RCS in dB of the solution I (half-RWG order) at unit directions."""
    out = []
    for s in np.atleast_2d(dirs):
        Es = np.zeros(3, dtype=complex)
        for u in range(C.n):
            vx, t = C.hrwg[u, 2], C.hrwg[u, 3]
            P = xyz[tri[t]]
            acc = np.zeros(3, dtype=complex)
            for g3 in range(3):
                g = R3[g3] @ P
                acc += (g - xyz[vx]) * (1 / 3) * np.exp(1j * k * (g @ s))
            Es += acc * (I[u] * C.length[u])
        es = 0.5 * np.cross(s, np.cross(s, Es))
        out.append(10 * np.log10(Z0 * Z0 * k * k * np.real(np.vdot(es, es)) / (4 * np.pi)))
    return np.array(out)


def expand(C, rwg):
    """n_hrwg x n_rwg map P with P x the half-RWG coefficients of RWG
    coefficients x on a conformal mesh: RWG e = half (v1, v2, vp, tp) minus
    half (v2, v1, vm, tm), the pair of edge number e (ContourEdge.cpp:26-34)."""
    P = np.zeros((C.n, len(rwg)))
    for u in range(C.n):
        e = C.hrwg[u, 4]
        P[u, e] = 1.0 if C.hrwg[u, 3] == rwg[e, 4] else -1.0
    return P


def current(xyz, tri, C, I):
    """Per-triangle current magnitudes of FastWT::calculateSurfaceCurrent
    (FastWT.cpp:1426-1472): the triangle's half-RWGs in list order (at most
    three); corner vertices, lengths and corner magnitudes carry over from
    the previous triangle where fewer than three are present (the reference
    declares them outside its loop).  Returns (nt x 4: magn1, magn2, magn3,
    magnc) and (nt x 9 corner vertices)."""
    nt = len(tri)
    per = [[] for _ in range(nt)]
    for u in range(C.n):
        per[C.hrwg[u, 3]].append(u)
    vx = np.zeros((3, 3))
    ln = np.zeros(3)
    mag = np.zeros(3)
    out, vout = np.zeros((nt, 4)), np.zeros((nt, 9))
    for t in range(nt):
        us = per[t][:3]
        for i, u in enumerate(us):
            vx[i] = xyz[C.hrwg[u, 2]]
            ln[i] = C.length[u]
        cen = vx.mean(0)
        la, lb, lc = (np.linalg.norm(vx[0] - vx[1]), np.linalg.norm(vx[1] - vx[2]), np.linalg.norm(vx[2] - vx[0]))
        sp = 0.5 * (la + lb + lc)
        dom = 2 * np.sqrt(max(0.0, sp * (sp - la) * (sp - lb) * (sp - lc)))
        cc = np.zeros(3, dtype=complex)
        for i, u in enumerate(us):
            cc += I[u] * (ln[i] / dom) * (cen - vx[i])
            nc = np.zeros(3, dtype=complex)
            for j, w in enumerate(us):
                if i != j:
                    nc += I[w] * (ln[j] / dom) * (vx[i] - vx[j])
            mag[i] = np.sqrt(np.sum(np.abs(nc) ** 2))
        out[t] = [mag[0], mag[1], mag[2], np.sqrt(np.sum(np.abs(cc) ** 2))]
        vout[t] = vx.ravel()
    return out, vout
