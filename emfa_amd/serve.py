"""ctypes binding of the solve server in libemfa_core.so (include/emfa_core.h).

A `Server` owns worker threads, each with its own device context; requests
are (mesh, frequency, formulation, precision, incidences) tuples, answered
with the monostatic RCS per incidence and, optionally, the surface current.
"""
import ctypes
import os
import threading
import warnings

import numpy as np

from . import hip

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_CORE = os.path.join(HERE, "libemfa_core.so")

EXPORTS = [
    "emfa_core_last_error", "emfa_core_config_get", "emfa_core_read_mesh", "emfa_core_run", "emfa_serve_create",
    "emfa_serve_destroy", "emfa_serve_submit", "emfa_serve_wait", "emfa_serve_fetch", "emfa_serve_drain",
    "emfa_serve_hold",
    "emfa_serve_get_stats", "emfa_serve_reset_stats", "emfa_serve_cached_factors", "emfa_serve_footprint",
    "emfa_core_write_formats",
]

ICOSPHERE, PLATE, USER = 0, 1, 2
DENSE, MLFMA, FASTWT = 0, 1, 2


class Request(ctypes.Structure):
    _fields_ = [("mesh_kind", ctypes.c_int), ("subdiv", ctypes.c_int), ("ax", ctypes.c_double),
                ("ay", ctypes.c_double), ("az", ctypes.c_double), ("formulation", ctypes.c_int),
                ("alpha", ctypes.c_double), ("freq", ctypes.c_double), ("precision", ctypes.c_int),
                ("quad", ctypes.c_int), ("nrhs", ctypes.c_int), ("angles", ctypes.POINTER(ctypes.c_double)),
                ("nv", ctypes.c_int), ("nt", ctypes.c_int), ("xyz", ctypes.POINTER(ctypes.c_double)),
                ("tri", ctypes.POINTER(ctypes.c_int)), ("solver", ctypes.c_int), ("mlfma_box", ctypes.c_double),
                ("iter_tol", ctypes.c_double), ("iter_max", ctypes.c_int), ("fastwt_beta", ctypes.c_double)]


class Result(ctypes.Structure):
    _fields_ = [("id", ctypes.c_longlong), ("status", ctypes.c_int), ("unknowns", ctypes.c_int),
                ("nrhs", ctypes.c_int), ("cache_hit", ctypes.c_int), ("refine_iterations", ctypes.c_int),
                ("relres", ctypes.c_double), ("fp64_fallback", ctypes.c_int), ("queue_ms", ctypes.c_double),
                ("service_ms", ctypes.c_double), ("latency_ms", ctypes.c_double), ("batch_size", ctypes.c_int)]


class Stats(ctypes.Structure):
    _fields_ = [("completed", ctypes.c_longlong), ("failed", ctypes.c_longlong), ("cache_hits", ctypes.c_longlong),
                ("p50_ms", ctypes.c_double), ("p99_ms", ctypes.c_double), ("mean_ms", ctypes.c_double),
                ("max_ms", ctypes.c_double), ("wall_s", ctypes.c_double), ("solves_per_s", ctypes.c_double),
                ("rhs_per_s", ctypes.c_double), ("peak_bytes", ctypes.c_size_t),
                ("pivot_fallbacks", ctypes.c_longlong), ("batched", ctypes.c_longlong),
                ("factor_batches", ctypes.c_longlong), ("batch_factors", ctypes.c_longlong)]


_CORE = None


def core():
    """Load libemfa_core.so (raises if it has not been built)."""
    global _CORE
    if _CORE is None:
        if not os.path.exists(LIB_CORE):
            raise hip.EmfaError(hip.EMFA_E_NODEVICE, f"{LIB_CORE} not built: run python -m emfa_amd.build")
        hip.lib()  # the core links against libemfa_hip.so
        L = ctypes.CDLL(LIB_CORE)
        L.emfa_core_last_error.restype = ctypes.c_char_p
        L.emfa_serve_submit.restype = ctypes.c_longlong
        L.emfa_serve_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_size_t, ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_void_p)]
        L.emfa_serve_wait.argtypes = [ctypes.c_void_p, ctypes.c_longlong, ctypes.POINTER(Result)]
        L.emfa_serve_fetch.argtypes = [ctypes.c_void_p, ctypes.c_longlong, ctypes.c_void_p, ctypes.c_int,
                                       ctypes.c_void_p, ctypes.c_longlong]
        L.emfa_serve_submit.argtypes = [ctypes.c_void_p, ctypes.POINTER(Request)]
        L.emfa_serve_footprint.restype = ctypes.c_size_t
        L.emfa_serve_footprint.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        for name in ("emfa_serve_destroy", "emfa_serve_drain", "emfa_serve_reset_stats",
                     "emfa_serve_cached_factors"):
            getattr(L, name).argtypes = [ctypes.c_void_p]
        L.emfa_serve_get_stats.argtypes = [ctypes.c_void_p, ctypes.POINTER(Stats)]
        L.emfa_serve_hold.argtypes = [ctypes.c_void_p, ctypes.c_int]
        _CORE = L
    return _CORE


def _err(code):
    return hip.EmfaError(code, core().emfa_core_last_error().decode())


def unknowns(kind, subdiv):
    """RWG count of a server mesh (closed icosphere: 1.5 T; plate: 3c^2 - 2c)."""
    if kind == PLATE:
        return 3 * subdiv * subdiv - 2 * subdiv
    return 30 * 4 ** subdiv


def footprint(n, nrhs, precision, factor_cached=False):
    return core().emfa_serve_footprint(int(n), int(nrhs), int(precision), int(bool(factor_cached)))


class Server:
    def __init__(self, device=0, workers=2, hbm_budget=64 << 30, keep_current=False):
        # Each worker drives two streams; HIP places streams on
        # GPU_MAX_HW_QUEUES hardware queues (4 unless set before HIP
        # initialises) and streams sharing a queue run in order.
        want = min(32, 2 * workers)
        try:
            have = int(os.environ.get("GPU_MAX_HW_QUEUES", "4"))
        except ValueError:
            have = 4
        if have < want:
            warnings.warn(f"Server with {workers} workers on {have} HIP hardware queues: the workers' streams "
                          f"serialise on them; set GPU_MAX_HW_QUEUES={want} before HIP initialises",
                          RuntimeWarning, stacklevel=2)
        self.h = ctypes.c_void_p()
        rc = core().emfa_serve_create(device, workers, int(hbm_budget), int(keep_current), ctypes.byref(self.h))
        if rc != 0:
            raise _err(hip.EMFA_E_NODEVICE if rc == 5 else hip.EMFA_E_INVALID)
        self.keep_current = keep_current
        self._nrhs = {}
        self._lock = threading.Lock()  # submit and wait may run on different threads

    def submit(self, angles, kind=ICOSPHERE, subdiv=2, axes=(1.0, 1.0, 1.0), formulation=hip.EFIE,
               alpha=1.0, freq=3e8, precision=hip.F64, quad=3, mesh=None, solver=DENSE, mlfma_box=0.25,
               iter_tol=1e-4, iter_max=600, fastwt_beta=0.0):
        """angles: (nrhs, 3) array of (theta, phi, pole) in degrees.  mesh:
        (xyz (nv, 3), tri (nt, 3) 0-based) of the caller's own mesh (kind
        becomes USER).  solver=MLFMA keeps the multipole operator resident and
        solves by GMRES(30) to iter_tol; solver=FASTWT does the same with the
        IP-SIE operator on the mesh's half-RWGs (CFIE; fastwt_beta 0: the
        default stabilization 0.1 |log10(0.07503)|), whose solutions are in
        half-RWG order."""
        a = np.ascontiguousarray(np.asarray(angles, dtype=np.float64).reshape(-1, 3))
        r = Request(kind, subdiv, axes[0], axes[1], axes[2], formulation, alpha, freq, precision, quad, a.shape[0],
                    a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        r.solver, r.mlfma_box, r.iter_tol, r.iter_max = solver, mlfma_box, iter_tol, iter_max
        r.fastwt_beta = fastwt_beta
        n = unknowns(kind, subdiv)
        if mesh is not None:
            xyz = np.ascontiguousarray(mesh[0], dtype=np.float64).reshape(-1, 3)
            tri = np.ascontiguousarray(mesh[1], dtype=np.int32).reshape(-1, 3)
            r.mesh_kind, r.nv, r.nt = USER, xyz.shape[0], tri.shape[0]
            r.xyz = xyz.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
            r.tri = tri.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
            n = hip.contour_build(xyz, tri).n if solver == FASTWT else hip.rwg_build(tri)[0].shape[0]
        with self._lock:
            rid = core().emfa_serve_submit(self.h, ctypes.byref(r))
            if rid < 0:
                raise _err(hip.EMFA_E_INVALID)
            self._nrhs[rid] = (a.shape[0], n)
        return rid

    def wait(self, rid=-1):
        """Result of request `rid` (any request if -1) as a dict, or None if nothing is pending."""
        res = Result()
        if core().emfa_serve_wait(self.h, rid, ctypes.byref(res)) != 0:
            return None
        err = core().emfa_core_last_error().decode() if res.status != 0 else ""
        with self._lock:
            nrhs, n = self._nrhs.pop(res.id)
        rcs = np.zeros(nrhs)
        cur = np.zeros(2 * n * nrhs) if (self.keep_current and res.status == 0) else None
        rc = core().emfa_serve_fetch(self.h, res.id, rcs.ctypes.data, nrhs,
                                     cur.ctypes.data if cur is not None else None, 0 if cur is None else cur.size)
        if rc != 0:
            raise _err(hip.EMFA_E_INVALID)
        out = {k: getattr(res, k) for k, _ in Result._fields_}
        out["error"] = err
        out["rcs_db"] = rcs
        if cur is not None:
            out["current"] = cur.view(np.complex128).reshape(n, nrhs)
        return out

    def drain(self):
        core().emfa_serve_drain(self.h)

    def hold(self, on=True):
        """While held, workers take no queued request; requests submitted
        together are then scheduled as one queue when the hold ends."""
        core().emfa_serve_hold(self.h, 1 if on else 0)

    def stats(self):
        s = Stats()
        core().emfa_serve_get_stats(self.h, ctypes.byref(s))
        return {k: getattr(s, k) for k, _ in Stats._fields_}

    def reset_stats(self):
        core().emfa_serve_reset_stats(self.h)

    def cached_factors(self):
        return core().emfa_serve_cached_factors(self.h)

    def close(self):
        if self.h:
            core().emfa_serve_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
