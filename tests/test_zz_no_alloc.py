"""emfa_bicgstab_noalloc (include/emfa_hip.h) against emfa_bicgstab through
the C ABI: the same bits in X and the same step counts, on two calls in a
row (the second on a smaller system), and no device memory taken from the
context's pool by the allocation-free entry point."""
import ctypes

import numpy as np
import pytest

from conftest import sphere
from emfa_amd import hip

pytestmark = pytest.mark.gpu


def _call(ctx, name, Z, B, X, tol, max_iter, variant):
    rep = hip.BicgstabReport()
    fn = getattr(hip.lib(), name)
    fn.restype = ctypes.c_int
    code = fn(ctx.h, Z.h, B.h, X.h, ctypes.c_double(tol), int(max_iter), int(variant), ctypes.byref(rep))
    return code, rep


def _system(ctx, subdiv, freq, angles, prec):
    xyz, tri, rwg = sphere(subdiv, 1.0)
    k = hip.wavenumber(freq)
    mesh = hip.Mesh(ctx, xyz, tri, rwg)
    Z = hip.Mat(ctx, mesh.n, mesh.n, prec)
    hip.fill(ctx, mesh, Z, hip.CFIE, k, 3, alpha=0.5)
    inc = hip.inc_table(angles)
    B = hip.Mat(ctx, mesh.n, inc.shape[0], hip.F64)
    hip.rhs_fill(ctx, mesh, B, hip.CFIE, k, 0.5, inc)
    return mesh, Z, B


def test_noalloc_bicgstab_matches_the_allocating_entry_point():
    lib = hip.lib()
    lib.emfa_ctx_bytes_in_use.restype = ctypes.c_size_t
    ctx = hip.Context(0)  # a fresh context: no scratch taken yet
    held = []
    try:
        # two calls in a row, the second on a smaller system
        for subdiv, freq, angles, prec in ((2, 6e8, [(0, 0, 0), (45, 30, 90), (120, 0, 0)], hip.F64),
                                           (1, 4e8, [(10, 0, 0), (60, 90, 0)], hip.F32)):
            print(f"no-alloc bicgstab: subdiv {subdiv}, {len(angles)} RHS, prec {prec}", flush=True)
            mesh, Z, B = _system(ctx, subdiv, freq, angles, prec)
            held += [mesh, Z, B]
            n, m = mesh.n, len(angles)
            for variant in (0, 1):
                X0 = hip.Mat(ctx, n, m, hip.F64)
                X1 = hip.Mat(ctx, n, m, hip.F64)
                held += [X0, X1]
                ctx.sync()
                before = lib.emfa_ctx_bytes_in_use(ctx.h)
                c0, r0 = _call(ctx, "emfa_bicgstab_noalloc", Z, B, X0, 1e-8, 500, variant)
                ctx.sync()
                assert lib.emfa_ctx_bytes_in_use(ctx.h) == before
                c1, r1 = _call(ctx, "emfa_bicgstab", Z, B, X1, 1e-8, 500, variant)
                assert c0 == c1 == hip.EMFA_OK, hip.lib().emfa_last_error()
                assert r0.iterations == r1.iterations and r0.max_steps == r1.max_steps
                assert r0.relres == r1.relres and r0.relres < 1e-7
                assert np.array_equal(X0.numpy(), X1.numpy())
    finally:
        for h in reversed(held):
            h.close()
        ctx.close()
