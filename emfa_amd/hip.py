"""ctypes binding of the C ABI in include/emfa_hip.h (libemfa_hip.so).

The product path: every call goes to the gfx950 library; there is no CPU
fallback.  If the library is missing, or no gfx950 device is present, the
calls raise.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libemfa_hip.so")

EMFA_OK, EMFA_E_INVALID, EMFA_E_OOM, EMFA_E_SINGULAR, EMFA_E_NOCONV, EMFA_E_HIP, EMFA_E_NODEVICE = range(7)
F32, F64 = 0, 1
EFIE, CFIE = 0, 1

# every symbol include/emfa_hip.h declares
EXPORTS = [
    "emfa_last_error", "emfa_version", "emfa_ctx_create", "emfa_ctx_destroy", "emfa_ctx_sync", "emfa_ctx_stream",
    "emfa_ctx_bytes_in_use", "emfa_ctx_cu_count", "emfa_ctx_set_cu_share", "emfa_ctx_set_panel_cap", "emfa_geom_icosphere", "emfa_geom_plate", "emfa_rwg_build", "emfa_contour_build", "emfa_mesh_upload",
    "emfa_mesh_destroy", "emfa_mesh_unknowns", "emfa_mat_create", "emfa_mat_destroy", "emfa_mat_upload",
    "emfa_mat_download", "emfa_mat_copy", "emfa_mat_rows", "emfa_mat_cols", "emfa_mat_prec", "emfa_mat_device_ptr",
    "emfa_fill", "emfa_rhs_fill", "emfa_lu_factor", "emfa_lu_factor_batch", "emfa_lu_destroy", "emfa_lu_solve", "emfa_lu_pivots",
    "emfa_lu_matrix", "emfa_lu_info", "emfa_refine", "emfa_gemm_residual", "emfa_gmres", "emfa_rcs",
    "emfa_prof_enable", "emfa_prof_get", "emfa_prof_reset", "emfa_mlfma_create", "emfa_mlfma_destroy",
    "emfa_mlfma_get_info", "emfa_mlfma_matvec", "emfa_mlfma_gmres", "emfa_lu_set_params", "emfa_lu_get_pivoting", "emfa_lu_autotune",
    "emfa_lu_set_update_mode", "emfa_bicgstab", "emfa_mlfma_bicgstab", "emfa_bicgstab_noalloc",
    "emfa_mlfma_boxes", "emfa_fmm_create", "emfa_fastwt_create", "emfa_fastwt_rhs", "emfa_fastwt_rcs", "emfa_ctx_set_graphs", "emfa_ctx_graph_stats",
    "emfa_mlfma_precondition",
]

_vp = ctypes.c_void_p
_d = ctypes.POINTER(ctypes.c_double)
_i = ctypes.POINTER(ctypes.c_int)


class FillParams(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("quad", ctypes.c_int), ("k", ctypes.c_double), ("alpha", ctypes.c_double),
                ("sigma_re", ctypes.c_double), ("sigma_im", ctypes.c_double), ("cfie_mode", ctypes.c_int)]


class RefineReport(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("relres_first", ctypes.c_double), ("relres_final", ctypes.c_double)]


class GmresReport(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("restarts", ctypes.c_int), ("relres", ctypes.c_double)]


class BicgstabReport(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("max_steps", ctypes.c_int), ("relres", ctypes.c_double)]


class LuTuning(ctypes.Structure):
    _fields_ = [("nb", ctypes.c_int), ("panel_variant", ctypes.c_int), ("panel_workgroups", ctypes.c_int),
                ("update_mode", ctypes.c_int), ("best_ms", ctypes.c_float), ("candidates", ctypes.c_int),
                ("cand_nb", ctypes.c_int * 16), ("cand_variant", ctypes.c_int * 16),
                ("cand_workgroups", ctypes.c_int * 16), ("cand_update_mode", ctypes.c_int * 16),
                ("cand_ms", ctypes.c_float * 16)]


class MlfmaParams(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("k", ctypes.c_double), ("alpha", ctypes.c_double),
                ("box_threshold", ctypes.c_double), ("cfie_mode", ctypes.c_int)]


class MlfmaInfo(ctypes.Structure):
    _fields_ = [("levels", ctypes.c_int), ("near_nnz", ctypes.c_longlong), ("bytes", ctypes.c_size_t),
                ("boxes", ctypes.c_int * 16), ("multipoles", ctypes.c_int * 16), ("box_edge", ctypes.c_double * 16),
                ("far_pairs", ctypes.c_longlong * 16), ("grid", ctypes.c_int * 3)]


class EmfaError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[emfa status {code}] {msg}")
        self.code = code


_LIB = None


def lib():
    """Load libemfa_hip.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise EmfaError(EMFA_E_NODEVICE, f"{LIB_PATH} not built: run python -m emfa_amd.build")
        L = ctypes.CDLL(LIB_PATH)
        L.emfa_last_error.restype = ctypes.c_char_p
        L.emfa_ctx_stream.restype = _vp
        L.emfa_ctx_bytes_in_use.restype = ctypes.c_size_t
        L.emfa_mat_device_ptr.restype = _vp
        L.emfa_lu_matrix.restype = _vp
        missing = [name for name in EXPORTS if not hasattr(L, name)]
        if missing:
            raise EmfaError(EMFA_E_NODEVICE, f"{LIB_PATH} lacks {', '.join(missing)}: rebuild it")
        _LIB = L
    return _LIB


def _check(code):
    if code != EMFA_OK:
        raise EmfaError(code, lib().emfa_last_error().decode())


def _pd(a):
    return a.ctypes.data_as(_d)


def _pi(a):
    return a.ctypes.data_as(_i)


# ------------------------------------------------------------------ geometry
def icosphere(subdiv, ax=1.0, ay=None, az=None):
    ay = ax if ay is None else ay
    az = ax if az is None else az
    nv, nt = ctypes.c_int(), ctypes.c_int()
    _check(lib().emfa_geom_icosphere(subdiv, ctypes.c_double(ax), ctypes.c_double(ay), ctypes.c_double(az), None, 0,
                                     None, 0, ctypes.byref(nv), ctypes.byref(nt)))
    xyz = np.zeros((nv.value, 3))
    tri = np.zeros((nt.value, 3), dtype=np.int32)
    _check(lib().emfa_geom_icosphere(subdiv, ctypes.c_double(ax), ctypes.c_double(ay), ctypes.c_double(az), _pd(xyz),
                                     nv.value, _pi(tri), nt.value, ctypes.byref(nv), ctypes.byref(nt)))
    return xyz, tri


def plate(cells, size=1.0):
    nv, nt = ctypes.c_int(), ctypes.c_int()
    _check(lib().emfa_geom_plate(cells, ctypes.c_double(size), None, 0, None, 0, ctypes.byref(nv), ctypes.byref(nt)))
    xyz = np.zeros((nv.value, 3))
    tri = np.zeros((nt.value, 3), dtype=np.int32)
    _check(lib().emfa_geom_plate(cells, ctypes.c_double(size), _pd(xyz), nv.value, _pi(tri), nt.value,
                                 ctypes.byref(nv), ctypes.byref(nt)))
    return xyz, tri


def rwg_build(tri):
    tri = np.ascontiguousarray(tri, dtype=np.int32)
    n, nb = ctypes.c_int(), ctypes.c_int()
    _check(lib().emfa_rwg_build(tri.shape[0], _pi(tri), None, 0, ctypes.byref(n), ctypes.byref(nb)))
    out = np.zeros((max(n.value, 1), 6), dtype=np.int32)
    _check(lib().emfa_rwg_build(tri.shape[0], _pi(tri), _pi(out), n.value, ctypes.byref(n), ctypes.byref(nb)))
    return out[:n.value].copy(), nb.value


class Contour:
    """Half-RWG list of the IP-SIE solver (emfa_contour_build; the
    reference's ContourEdge with policy::HALF_RWG).  hrwg: n x 6 (v1, v2, vx,
    triangle, edge number, type 0/1), length: n, and the contour lines of
    the nonconformal half-RWGs as CSR (line_ptr, line_tri, line_seg n_lines x 6)."""

    def __init__(self, hrwg, length, line_ptr, line_tri, line_seg):
        self.hrwg, self.length = hrwg, length
        self.line_ptr, self.line_tri, self.line_seg = line_ptr, line_tri, line_seg

    @property
    def n(self):
        return self.hrwg.shape[0]

    def line_length(self, i):
        """FP32 lengths of the contour lines of nonconformal edge i."""
        seg = self.line_seg[self.line_ptr[i]:self.line_ptr[i + 1]].astype(np.float32)
        d = seg[:, 3:] - seg[:, :3]
        return np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2])


def contour_build(xyz, tri):
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    tri = np.ascontiguousarray(tri, dtype=np.int32)
    n, nnc, nl = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    L = lib()
    _check(L.emfa_contour_build(xyz.shape[0], _pd(xyz), tri.shape[0], _pi(tri), None, None, 0, ctypes.byref(n),
                                None, None, None, 0, ctypes.byref(nnc), ctypes.byref(nl)))
    h = np.zeros((max(n.value, 1), 6), dtype=np.int32)
    ln = np.zeros(max(n.value, 1))
    ptr = np.zeros(nnc.value + 1, dtype=np.int32)
    lt = np.zeros(max(nl.value, 1), dtype=np.int32)
    seg = np.zeros((max(nl.value, 1), 6))
    _check(L.emfa_contour_build(xyz.shape[0], _pd(xyz), tri.shape[0], _pi(tri), _pi(h), _pd(ln), h.shape[0],
                                ctypes.byref(n), _pi(ptr), _pi(lt), _pd(seg), lt.shape[0], ctypes.byref(nnc),
                                ctypes.byref(nl)))
    return Contour(h[:n.value].copy(), ln[:n.value].copy(), ptr, lt[:nl.value].copy(), seg[:nl.value].copy())


def wavenumber(freq):
    """k = 2 pi f / c0 (EMSolver/EFIE.cpp:47)."""
    return 2.0 * np.pi * float(freq) / 299792458.0


def incidence(theta_deg, phi_deg, pol_deg):
    """(k_hat, e_hat, h_hat) as EMSolver/EM.cpp:29-41 builds them."""
    th, ph, po = np.radians([theta_deg, phi_deg, pol_deg])
    k = np.array([-np.sin(th) * np.cos(ph), -np.sin(th) * np.sin(ph), -np.cos(th)])
    e = np.array([-np.cos(po) * np.cos(th) * np.cos(ph) - np.sin(po) * np.sin(ph),
                  -np.cos(po) * np.cos(th) * np.sin(ph) + np.sin(po) * np.cos(ph), np.cos(po) * np.sin(th)])
    return k, e, np.cross(k, e)


def inc_table(angles):
    return np.array([np.concatenate(incidence(*a)) for a in angles]).reshape(-1, 9)


# ------------------------------------------------------------------ handles
class Context:
    def __init__(self, device=0):
        self.h = _vp()
        _check(lib().emfa_ctx_create(device, ctypes.byref(self.h)))

    def close(self):
        if self.h:
            lib().emfa_ctx_destroy(self.h)
            self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        _check(lib().emfa_ctx_sync(self.h))

    def stream(self):
        return lib().emfa_ctx_stream(self.h)

    def bytes_in_use(self):
        return lib().emfa_ctx_bytes_in_use(self.h)

    def cu_count(self):
        return lib().emfa_ctx_cu_count(self.h)

    def lu_params(self, nb=0, panel_workgroups=0, panel_variant=-1):
        _check(lib().emfa_lu_set_params(self.h, int(nb), int(panel_workgroups), int(panel_variant)))

    def lu_update_mode(self, mode):
        """0: FP32 MFMA trailing update; 1: three-term bf16 split on bf16 MFMA."""
        _check(lib().emfa_lu_set_update_mode(self.h, int(mode)))

    def autotune_lu(self, n, prec, for_refinement=False):
        """Time the LU candidates at size n, keep the fastest; returns the report.
        for_refinement admits the FP32 variants meant for FP64-refined solves
        (bf16-split update, panels without interchanges)."""
        t = LuTuning()
        _check(lib().emfa_lu_autotune(self.h, int(n), int(prec), int(bool(for_refinement)), ctypes.byref(t)))
        c = t.candidates
        return {"nb": t.nb, "panel_variant": t.panel_variant, "panel_workgroups": t.panel_workgroups,
                "update": "bf16x3" if t.update_mode == 1 else "fp32", "best_ms": t.best_ms,
                "candidates": [{"nb": t.cand_nb[i], "panel_variant": t.cand_variant[i],
                                "panel_workgroups": t.cand_workgroups[i],
                                "update": "bf16x3" if t.cand_update_mode[i] == 1 else "fp32",
                                "ms": round(t.cand_ms[i], 2)}
                               for i in range(min(c, 16))]}

    def set_cu_share(self, cus):
        _check(lib().emfa_ctx_set_cu_share(self.h, int(cus)))

    def graphs(self, on=True):
        """Run pivot-free bf16-split FP32 factorisations as captured HIP graphs."""
        _check(lib().emfa_ctx_set_graphs(self.h, int(bool(on))))

    def graph_stats(self):
        """(graphs captured, graph launches) of this context."""
        a, b = ctypes.c_longlong(), ctypes.c_longlong()
        _check(lib().emfa_ctx_graph_stats(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def prof(self, on=True):
        _check(lib().emfa_prof_enable(self.h, int(on)))

    def prof_reset(self):
        _check(lib().emfa_prof_reset(self.h))

    def prof_get(self, family):
        ms, n, w = ctypes.c_double(), ctypes.c_longlong(), ctypes.c_double()
        _check(lib().emfa_prof_get(self.h, family.encode(), ctypes.byref(ms), ctypes.byref(n), ctypes.byref(w)))
        return ms.value, n.value, w.value


class Mesh:
    def __init__(self, ctx, xyz, tri, rwg):
        self.ctx = ctx
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        tri = np.ascontiguousarray(tri, dtype=np.int32)
        rwg = np.ascontiguousarray(rwg, dtype=np.int32)
        self.xyz, self.tri, self.rwg = xyz, tri, rwg
        self.h = _vp()
        _check(lib().emfa_mesh_upload(ctx.h, _pd(xyz), xyz.shape[0], _pi(tri), tri.shape[0], _pi(rwg), rwg.shape[0],
                                      ctypes.byref(self.h)))
        self.n = rwg.shape[0]

    def close(self):
        if self.h and self.ctx.h:
            lib().emfa_mesh_destroy(self.h)
        self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Mat:
    def __init__(self, ctx, rows, cols, prec=F64, handle=None, owned=True):
        self.ctx = ctx
        self.rows, self.cols, self.prec = rows, cols, prec
        self.owned = owned
        if handle is None:
            self.h = _vp()
            _check(lib().emfa_mat_create(ctx.h, rows, cols, prec, ctypes.byref(self.h)))
        else:
            self.h = _vp(handle)

    @classmethod
    def from_numpy(cls, ctx, a, prec=F64):
        a = np.ascontiguousarray(a, dtype=np.complex128)
        if a.ndim == 1:
            a = a.reshape(-1, 1)
        m = cls(ctx, a.shape[0], a.shape[1], prec)
        m.upload(a)
        return m

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.complex128).reshape(self.rows, self.cols)
        _check(lib().emfa_mat_upload(self.h, a.ctypes.data_as(_d)))

    def numpy(self):
        out = np.zeros((self.rows, self.cols), dtype=np.complex128)
        _check(lib().emfa_mat_download(self.h, out.ctypes.data_as(_d)))
        return out

    def copy_from(self, other):
        _check(lib().emfa_mat_copy(self.h, other.h))

    def close(self):
        if self.h and self.owned and self.ctx.h:
            lib().emfa_mat_destroy(self.h)
        self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class LU:
    def __init__(self, ctx, Z, prec, handle=None):
        self.ctx = ctx
        self.n = Z.rows
        self.prec = prec
        if handle is not None:  # a factor made by lu_factor_batch
            self.h = handle
            self.singular = lib().emfa_lu_info(handle) != 0
            return
        self.h = _vp()
        code = lib().emfa_lu_factor(ctx.h, Z.h, prec, ctypes.byref(self.h))
        self.singular = code == EMFA_E_SINGULAR
        if code not in (EMFA_OK, EMFA_E_SINGULAR):
            _check(code)

    def info(self):
        return lib().emfa_lu_info(self.h)

    def pivots(self):
        p = np.zeros(self.n, dtype=np.int32)
        _check(lib().emfa_lu_pivots(self.h, _pi(p)))
        return p

    def pivoting(self):
        """(interchanges, fallback): 1/0 partial pivoting or none; 1 when a
        factorisation without interchanges grew and was redone with them."""
        a, b = ctypes.c_int(), ctypes.c_int()
        _check(lib().emfa_lu_get_pivoting(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def factor(self):
        return Mat(self.ctx, self.n, self.n, self.prec, handle=lib().emfa_lu_matrix(self.h), owned=False)

    def solve(self, B):
        _check(lib().emfa_lu_solve(self.ctx.h, self.h, B.h))

    def close(self):
        if self.h and self.ctx.h:
            lib().emfa_lu_destroy(self.h)
        self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def lu_factor_batch(ctx, mats, prec):
    """Factor same-size matrices with one launch sequence where the context
    allows it (emfa_lu_factor_batch); returns one LU per matrix."""
    mats = list(mats)
    arr = (_vp * len(mats))(*[m.h for m in mats])
    outs = (_vp * len(mats))()
    code = lib().emfa_lu_factor_batch(ctx.h, arr, len(mats), prec, outs)
    if code not in (EMFA_OK, EMFA_E_SINGULAR):
        _check(code)
    return [LU(ctx, m, prec, handle=_vp(h)) for m, h in zip(mats, outs)]


# ------------------------------------------------------------------ operations
def fill(ctx, mesh, Z, kind, k, quad=3, alpha=1.0, sigma=0j, cfie_mode=0):
    p = FillParams(kind, quad, float(k), float(alpha), float(np.real(sigma)), float(np.imag(sigma)), int(cfie_mode))
    _check(lib().emfa_fill(ctx.h, mesh.h, ctypes.byref(p), Z.h))


def rhs_fill(ctx, mesh, B, kind, k, alpha, inc):
    inc = np.ascontiguousarray(inc, dtype=np.float64).reshape(-1, 9)
    _check(lib().emfa_rhs_fill(ctx.h, mesh.h, kind, ctypes.c_double(k), ctypes.c_double(alpha), inc.shape[0],
                               _pd(inc), B.h))


def refine(ctx, Z, lu, B, X, tol=1e-12, maxit=10):
    rep = RefineReport()
    code = lib().emfa_refine(ctx.h, Z.h, lu.h, B.h, X.h, ctypes.c_double(tol), maxit, ctypes.byref(rep))
    if code not in (EMFA_OK, EMFA_E_NOCONV):
        _check(code)
    return rep, code == EMFA_OK


def gemm_residual(ctx, A, X, C, negate=True, accumulate=True):
    _check(lib().emfa_gemm_residual(ctx.h, A.h, X.h, C.h, int(negate), int(accumulate)))


def bicgstab(ctx, Z, B, X, tol=1e-2, max_iter=500, variant=0):
    """BiCGSTAB from x0 = 0 (variant 0: EMIterative/iml.cpp:3-39, as CFIE.cpp:76
    calls it; variant 1: the matrix-free overload iml.cpp:88-154).
    Returns (report, converged)."""
    rep = BicgstabReport()
    code = lib().emfa_bicgstab(ctx.h, Z.h, B.h, X.h, ctypes.c_double(tol), int(max_iter), int(variant),
                               ctypes.byref(rep))
    if code not in (EMFA_OK, EMFA_E_NOCONV):
        _check(code)
    return rep, code == EMFA_OK


def gmres(ctx, Z, B, X, m=30, tol=1e-6, max_restart=100):
    rep = GmresReport()
    code = lib().emfa_gmres(ctx.h, Z.h, B.h, X.h, m, ctypes.c_double(tol), max_restart, ctypes.byref(rep))
    if code not in (EMFA_OK, EMFA_E_NOCONV):
        _check(code)
    return rep, code == EMFA_OK


class MLFMA:
    """Multilevel fast multipole operator of a mesh (emfa_mlfma_*)."""

    _create = "emfa_mlfma_create"

    def __init__(self, ctx, mesh, k, alpha=0.5, box_threshold=0.25, kind=CFIE, cfie_mode=0):
        self.ctx = ctx
        self.h = _vp()
        p = MlfmaParams(kind, float(k), float(alpha), float(box_threshold), int(cfie_mode))
        _check(getattr(lib(), self._create)(ctx.h, mesh.h, ctypes.byref(p), ctypes.byref(self.h)))
        self.n = mesh.rwg.shape[0]

    def info(self):
        i = MlfmaInfo()
        _check(lib().emfa_mlfma_get_info(self.h, ctypes.byref(i)))
        lv = i.levels
        return {"levels": lv, "near_nnz": i.near_nnz, "bytes": i.bytes, "boxes": list(i.boxes[:lv + 1]),
                "multipoles": list(i.multipoles[:lv + 1]), "box_edge": list(i.box_edge[:lv + 1]),
                "far_pairs": list(i.far_pairs[:lv + 1]), "grid": list(i.grid)}

    def boxes(self, level):
        """(near boxes, far boxes, unknowns) per box of a level, box order."""
        nb = self.info()["boxes"][level]
        a, b, c = (np.zeros(nb, dtype=np.int32) for _ in range(3))
        _check(lib().emfa_mlfma_boxes(self.h, int(level), a.ctypes.data_as(_i), b.ctypes.data_as(_i),
                                      c.ctypes.data_as(_i)))
        return a, b, c

    def matvec(self, X, Y):
        _check(lib().emfa_mlfma_matvec(self.ctx.h, self.h, X.h, Y.h))

    def precondition(self, mesh, row_threshold=0.9, col_threshold=0.87):
        """Sparse approximate inverse (emfa_mlfma_precondition); the operator
        becomes P Z and the solvers precondition their right-hand sides.
        Returns P's nonzero count."""
        e = ctypes.c_longlong()
        _check(lib().emfa_mlfma_precondition(self.ctx.h, self.h, mesh.h, ctypes.c_double(row_threshold),
                                             ctypes.c_double(col_threshold), ctypes.byref(e)))
        return e.value

    def bicgstab(self, B, X, tol=1e-4, max_iter=500, variant=0):
        rep = BicgstabReport()
        code = lib().emfa_mlfma_bicgstab(self.ctx.h, self.h, B.h, X.h, ctypes.c_double(tol), int(max_iter),
                                         int(variant), ctypes.byref(rep))
        if code not in (EMFA_OK, EMFA_E_NOCONV):
            _check(code)
        return rep, code == EMFA_OK

    def gmres(self, B, X, m=30, tol=1e-4, max_restart=50):
        rep = GmresReport()
        code = lib().emfa_mlfma_gmres(self.ctx.h, self.h, B.h, X.h, m, ctypes.c_double(tol), max_restart,
                                      ctypes.byref(rep))
        if code not in (EMFA_OK, EMFA_E_NOCONV):
            _check(code)
        return rep, code == EMFA_OK

    def close(self):
        if self.h and self.ctx.h:
            lib().emfa_mlfma_destroy(self.h)
        self.h = _vp()


class FMM(MLFMA):
    """Single-level fast multipole operator (emfa_fmm_create; box_threshold is
    the box length in wavelengths, `fmm_box_length:`)."""
    _create = "emfa_fmm_create"

    def __init__(self, ctx, mesh, k, alpha=0.5, box_length=0.4, kind=CFIE, cfie_mode=0):
        super().__init__(ctx, mesh, k, alpha, box_length, kind, cfie_mode)


class FastWT(MLFMA):
    """IP-SIE operator on half-RWGs (emfa_fastwt_create); `contour` is
    contour_build's result for the mesh, beta the stabilization factor times
    |log10(average size)|.  Vectors are in half-RWG order."""

    def __init__(self, ctx, mesh, contour, k, alpha=0.5, box_threshold=0.25, beta=0.1 * 1.1248, cfie_mode=0):
        self.ctx = ctx
        self.h = _vp()
        self.mesh = mesh
        c = contour
        p = MlfmaParams(CFIE, float(k), float(alpha), float(box_threshold), int(cfie_mode))
        h = np.ascontiguousarray(c.hrwg, dtype=np.int32)
        ln = np.ascontiguousarray(c.length, dtype=np.float64)
        ptr = np.ascontiguousarray(c.line_ptr, dtype=np.int32)
        lt = np.ascontiguousarray(c.line_tri, dtype=np.int32)
        seg = np.ascontiguousarray(c.line_seg, dtype=np.float64)
        nnc = len(ptr) - 1
        _check(lib().emfa_fastwt_create(ctx.h, mesh.h, h.shape[0], _pi(h), _pd(ln), nnc, _pi(ptr) if nnc else None,
                                        _pi(lt) if nnc else None, _pd(seg) if nnc else None, ctypes.byref(p),
                                        ctypes.c_double(beta), ctypes.byref(self.h)))
        self.n = h.shape[0]

    def rhs(self, B, inc):
        inc = np.ascontiguousarray(inc, dtype=np.float64).reshape(-1, 9)
        _check(lib().emfa_fastwt_rhs(self.ctx.h, self.h, self.mesh.h, B.h, inc.shape[0], _pd(inc)))

    def rcs(self, X, dirs, col=0):
        dirs = np.ascontiguousarray(dirs, dtype=np.float64).reshape(-1, 3)
        out = np.zeros(dirs.shape[0])
        _check(lib().emfa_fastwt_rcs(self.ctx.h, self.h, self.mesh.h, X.h, int(col), dirs.shape[0], _pd(dirs),
                                     _pd(out)))
        return out


def rcs(ctx, mesh, k, X, dirs, cols=None):
    dirs = np.ascontiguousarray(dirs, dtype=np.float64).reshape(-1, 3)
    out = np.zeros(dirs.shape[0])
    c = None
    if cols is not None:
        c = np.ascontiguousarray(cols, dtype=np.int32)
    _check(lib().emfa_rcs(ctx.h, mesh.h, ctypes.c_double(k), X.h, dirs.shape[0], _pd(dirs),
                          _pi(c) if c is not None else None, _pd(out)))
    return out
