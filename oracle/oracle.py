"""ctypes front end of the FP64 CPU oracle (oracle/emfa_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.  Each wrapper
names the reference routine its C function restates.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_d = ctypes.POINTER(ctypes.c_double)
_i = ctypes.POINTER(ctypes.c_int)


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.orc_build_rwg.restype = ctypes.c_int
        L.orc_build_rwg.argtypes = [ctypes.c_int, _i, _i, ctypes.c_int, _i]
        L.orc_incidence.argtypes = [ctypes.c_double] * 3 + [_d] * 3
        L.orc_fill_efie.argtypes = [_d, _i, ctypes.c_int, _i, ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_int, ctypes.c_int, _d]
        L.orc_fill_cfie.argtypes = [_d, _i, ctypes.c_int, _i, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_int, ctypes.c_int, _d]
        L.orc_fill_hrwg.argtypes = [_d, _i, ctypes.c_int, _i, _d, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_int, ctypes.c_int, _d]
        L.orc_rhs.argtypes = [_d, _i, ctypes.c_int, _i, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                              ctypes.c_int, _d, _d]
        L.orc_rcs.argtypes = [_d, _i, ctypes.c_int, _i, ctypes.c_double, _d, ctypes.c_int, ctypes.c_int, _d, _i, _d]
        L.orc_lu.restype = ctypes.c_int
        L.orc_lu.argtypes = [ctypes.c_int, _d, _i]
        L.orc_lu_solve.argtypes = [ctypes.c_int, _d, _i, ctypes.c_int, _d]
        L.orc_gmres.restype = ctypes.c_int
        L.orc_gmres.argtypes = [ctypes.c_int, _d, _d, _d, ctypes.c_int, ctypes.c_double, ctypes.c_int, _d]
        L.orc_bicgstab.restype = ctypes.c_int
        L.orc_bicgstab.argtypes = [ctypes.c_int, _d, _d, _d, ctypes.c_double, ctypes.c_int, _i]
        L.orc_bicgstab_mf.restype = ctypes.c_int
        L.orc_bicgstab_mf.argtypes = [ctypes.c_int, _d, _d, _d, ctypes.c_double, ctypes.c_int, _d]
        L.orc_current.argtypes = [_d, ctypes.c_int, ctypes.c_int, _i, _d, ctypes.c_int, ctypes.c_int, _d, _d]
        L.orc_quad_rule.restype = ctypes.c_int
        L.orc_quad_rule.argtypes = [ctypes.c_int, _d, _d]
        L.orc_wavenumber.restype = ctypes.c_double
        L.orc_wavenumber.argtypes = [ctypes.c_double]
        L.orc_potentials.argtypes = [_d, _d, _d, _d]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_max_threads.restype = ctypes.c_int
        _LIB = L
    return _LIB


def _pd(a):
    return a.ctypes.data_as(_d)


def _pi(a):
    return a.ctypes.data_as(_i)


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def set_threads(n):
    lib().orc_set_threads(int(n))


def max_threads():
    return lib().orc_max_threads()


def wavenumber(freq):
    """k = 2 pi f / c (EMSolver/EFIE.cpp:47)."""
    return lib().orc_wavenumber(float(freq))


def quad_rule(q):
    """Triangle rule (EMComponent/Quadrature.h:20-27, 42-72): barycentrics and weights."""
    bary = np.zeros((q, 3))
    w = np.zeros(q)
    lib().orc_quad_rule(q, _pd(bary), _pd(w))
    return bary, w


def build_rwg(tri):
    """RWG list by the reference's O(T*E) scan (EMComponent/CommonEdge.cpp:74-106,704-743).

    Returns (rwg[N,6] int32 {v1,v2,vxp,vxm,tp,tm}, n_boundary_edges)."""
    tri = _c(tri, np.int32)
    nt = tri.shape[0]
    out = np.zeros((3 * nt, 6), dtype=np.int32)
    nb = ctypes.c_int(0)
    n = lib().orc_build_rwg(nt, _pi(tri), _pi(out), 3 * nt, ctypes.byref(nb))
    if n < 0:
        raise RuntimeError("rwg capacity")
    return out[:n].copy(), nb.value


def incidence(theta_deg, phi_deg, pol_deg):
    """(k_hat, e_hat, h_hat) of EMSolver/EM.cpp:29-41."""
    k = np.zeros(3)
    e = np.zeros(3)
    h = np.zeros(3)
    lib().orc_incidence(float(theta_deg), float(phi_deg), float(pol_deg), _pd(k), _pd(e), _pd(h))
    return k, e, h


def inc_table(angles):
    """angles: iterable of (theta, phi, pol) degrees -> (nrhs, 9) {k, e, h}."""
    rows = []
    for th, ph, po in angles:
        k, e, h = incidence(th, ph, po)
        rows.append(np.concatenate([k, e, h]))
    return np.array(rows, dtype=np.float64).reshape(-1, 9)


def fill_efie(xyz, tri, rwg, k, quad=3, sigma=0j, rows=None):
    """EFIE impedance matrix, EMSolver/EFIE.cpp:603-696 (row-major complex128)."""
    xyz = _c(xyz, np.float64)
    tri = _c(tri, np.int32)
    rwg = _c(rwg, np.int32)
    n = rwg.shape[0]
    r0, r1 = (0, n) if rows is None else rows
    Z = np.zeros((r1 - r0, n), dtype=np.complex128)
    lib().orc_fill_efie(_pd(xyz), _pi(tri), n, _pi(rwg), float(k), int(quad), float(np.real(sigma)),
                        float(np.imag(sigma)), r0, r1, Z.ctypes.data_as(_d))
    return Z


def fill_cfie(xyz, tri, rwg, k, alpha, quad=3, mode=0, rows=None):
    """CFIE impedance matrix, EMSolver/CFIE.cpp:485-568 + Miscellaneous.cpp:6-77."""
    xyz = _c(xyz, np.float64)
    tri = _c(tri, np.int32)
    rwg = _c(rwg, np.int32)
    n = rwg.shape[0]
    r0, r1 = (0, n) if rows is None else rows
    Z = np.zeros((r1 - r0, n), dtype=np.complex128)
    lib().orc_fill_cfie(_pd(xyz), _pi(tri), n, _pi(rwg), float(k), float(alpha), int(quad), int(mode), r0, r1,
                        Z.ctypes.data_as(_d))
    return Z


def fill_hrwg(xyz, tri, hrwg, length, k, alpha, quad=3, mode=0, rows=None):
    """Surface part l_f l_s K(f, s) of the IP-SIE near entries between
    half-RWGs (EMSolver/FastWT.cpp:753-762 with Miscellaneous.cpp:6-77)."""
    xyz = _c(xyz, np.float64)
    tri = _c(tri, np.int32)
    hrwg = _c(hrwg, np.int32)
    length = _c(length, np.float64)
    n = hrwg.shape[0]
    r0, r1 = (0, n) if rows is None else rows
    Z = np.zeros((r1 - r0, n), dtype=np.complex128)
    lib().orc_fill_hrwg(_pd(xyz), _pi(tri), n, _pi(hrwg), _pd(length), float(k), float(alpha), int(quad), int(mode),
                        r0, r1, Z.ctypes.data_as(_d))
    return Z


def rhs(xyz, tri, rwg, kind, k, alpha, inc):
    """Excitation columns: EFIE EFIE.cpp:568-601,698-722; CFIE Miscellaneous.cpp:152-189."""
    xyz = _c(xyz, np.float64)
    tri = _c(tri, np.int32)
    rwg = _c(rwg, np.int32)
    inc = _c(inc, np.float64).reshape(-1, 9)
    n = rwg.shape[0]
    B = np.zeros((n, inc.shape[0]), dtype=np.complex128)
    lib().orc_rhs(_pd(xyz), _pi(tri), n, _pi(rwg), int(kind), float(k), float(alpha), inc.shape[0], _pd(inc),
                  B.ctypes.data_as(_d))
    return B


def rcs(xyz, tri, rwg, k, X, dirs, cols=None):
    """RCS in dB per direction, EMSolver/EFIE.cpp:788-832."""
    xyz = _c(xyz, np.float64)
    tri = _c(tri, np.int32)
    rwg = _c(rwg, np.int32)
    X = _c(X, np.complex128)
    if X.ndim == 1:
        X = X.reshape(-1, 1)
    dirs = _c(dirs, np.float64).reshape(-1, 3)
    nd = dirs.shape[0]
    c = None
    if cols is not None:
        c = _c(cols, np.int32)
    out = np.zeros(nd)
    lib().orc_rcs(_pd(xyz), _pi(tri), rwg.shape[0], _pi(rwg), float(k), X.ctypes.data_as(_d), X.shape[1], nd,
                  _pd(dirs), _pi(c) if c is not None else None, _pd(out))
    return out


def lu(A):
    """Partial-pivot LU, LAPACK ?getrf semantics (EMSolver/MFIE.cpp:64 -> cgesv). Row-major."""
    A = np.array(A, dtype=np.complex128, order="C", copy=True)
    n = A.shape[0]
    piv = np.zeros(n, dtype=np.int32)
    info = lib().orc_lu(n, A.ctypes.data_as(_d), _pi(piv))
    return A, piv, info


def lu_solve(LU, piv, B):
    LU = _c(LU, np.complex128)
    B = np.array(B, dtype=np.complex128, order="C", copy=True)
    if B.ndim == 1:
        B = B.reshape(-1, 1)
    lib().orc_lu_solve(LU.shape[0], LU.ctypes.data_as(_d), _pi(_c(piv, np.int32)), B.shape[1],
                       B.ctypes.data_as(_d))
    return B


def gmres(A, b, m=30, tol=1e-6, max_restart=100):
    """Restarted GMRES (EMIterative/iml.cpp:156-207 with its defects corrected)."""
    A = _c(A, np.complex128)
    b = _c(b, np.complex128).ravel()
    x = np.zeros_like(b)
    rr = ctypes.c_double(0)
    it = lib().orc_gmres(A.shape[0], A.ctypes.data_as(_d), b.ctypes.data_as(_d), x.ctypes.data_as(_d), int(m),
                         float(tol), int(max_restart), ctypes.byref(rr))
    return x, it, rr.value


def bicgstab(A, b, eps=1e-2, max_iter=500):
    """BiCGSTAB of EMIterative/iml.cpp:3-39."""
    A = _c(A, np.complex128)
    b = _c(b, np.complex128).ravel()
    x = np.zeros_like(b)
    ok = ctypes.c_int(0)
    it = lib().orc_bicgstab(A.shape[0], A.ctypes.data_as(_d), b.ctypes.data_as(_d), x.ctypes.data_as(_d),
                            float(eps), int(max_iter), ctypes.byref(ok))
    return x, it, bool(ok.value)


def bicgstab_mf(A, b, eps=1e-2, max_iter=500):
    """BiCGSTAB of EMIterative/iml.cpp:88-154 (the matrix-free overload, with
    its shadow-residual restart).  Returns (x, steps, relres)."""
    A = _c(A, np.complex128)
    b = _c(b, np.complex128).ravel()
    x = np.zeros_like(b)
    rr = ctypes.c_double(0)
    it = lib().orc_bicgstab_mf(A.shape[0], A.ctypes.data_as(_d), b.ctypes.data_as(_d), x.ctypes.data_as(_d),
                               float(eps), int(max_iter), ctypes.byref(rr))
    return x, it, rr.value


def current(xyz, nt, rwg, X, col=0):
    """Per-triangle current magnitudes (EMSolver/CFIE.cpp:727-794)."""
    xyz = _c(xyz, np.float64)
    rwg = _c(rwg, np.int32)
    X = _c(X, np.complex128)
    if X.ndim == 1:
        X = X.reshape(-1, 1)
    out = np.zeros((nt, 4))
    vout = np.zeros((nt, 9))
    lib().orc_current(_pd(xyz), nt, rwg.shape[0], _pi(rwg), X.ctypes.data_as(_d), X.shape[1], int(col), _pd(out),
                      _pd(vout))
    return out, vout


def potentials(tri9, r):
    """Analytic (Is1, Is2) of EMSolver/EFIE.cpp:259-291 for triangle tri9 (3x3) seen from r."""
    t = _c(tri9, np.float64).reshape(9)
    rr = _c(r, np.float64).reshape(3)
    is1 = np.zeros(3)
    is2 = np.zeros(1)
    lib().orc_potentials(_pd(t), _pd(rr), _pd(is1), _pd(is2))
    return is1, float(is2[0])
