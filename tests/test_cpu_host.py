"""Host-side checks that need no GPU: the C-ABI libraries load and export
every declared symbol, host geometry and RWG ordering, the configuration
loader's semantics, the NAS reader, and the launcher's error paths."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import meshgen
from conftest import ROOT
from emfa_amd import hip
from oracle import oracle as O

CORE = os.path.join(ROOT, "emfa_amd", "libemfa_core.so")
EXE = os.path.join(ROOT, "emfa_amd", "bin", "emfa")


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(emfa_[a-z0-9_]+)\s*\(", text))


def test_hip_library_exports_every_declared_symbol():
    L = hip.lib()
    declared = _declared("emfa_hip.h")
    assert declared == set(hip.EXPORTS)
    for name in declared:
        assert hasattr(L, name), name
    nm = subprocess.run(["nm", "-D", "--defined-only", hip.LIB_PATH], stdout=subprocess.PIPE, text=True).stdout
    exported = set(re.findall(r" T (emfa_[a-z0-9_]+)", nm))
    assert declared <= exported


def test_core_library_exports_every_declared_symbol():
    from emfa_amd import serve
    L = serve.core()
    declared = _declared("emfa_core.h")
    assert declared == set(serve.EXPORTS)
    for name in declared:
        assert hasattr(L, name), name
    nm = subprocess.run(["nm", "-D", "--defined-only", CORE], stdout=subprocess.PIPE, text=True).stdout
    assert declared <= set(re.findall(r" T (emfa_[a-z0-9_]+)", nm))


def test_server_footprint_model():
    from emfa_amd import serve
    n, m = 1920, 16
    cached = serve.footprint(n, m, hip.F32, True)
    assert serve.footprint(n, m, hip.F32) - cached >= 2 * n * n * 8  # Z and its factor
    assert serve.footprint(n, m, hip.F64) - serve.footprint(n, m, hip.F64, True) >= 2 * n * n * 16
    assert serve.unknowns(serve.PLATE, 14) == 560 and serve.unknowns(serve.ICOSPHERE, 3) == 1920


def _has_gpu():
    return os.path.exists("/dev/kfd")


@pytest.mark.skipif(_has_gpu(), reason="checks the behaviour without a device")
def test_context_fails_loudly_without_device():
    with pytest.raises(hip.EmfaError) as e:
        hip.Context(0)
    assert e.value.code == hip.EMFA_E_NODEVICE


# ---------------------------------------------------------------- geometry
@pytest.mark.parametrize("s", [0, 1, 2, 3, 4])
def test_icosphere_matches_independent_generator(s):
    xyz, tri = hip.icosphere(s, 0.7)
    x2, t2 = meshgen.icosphere(s, 0.7)
    assert tri.shape == (20 * 4 ** s, 3) and xyz.shape == (10 * 4 ** s + 2, 3)
    assert np.array_equal(tri, t2)
    assert np.abs(xyz - x2).max() < 1e-15
    assert np.allclose(np.linalg.norm(xyz, axis=1), 0.7, atol=1e-14)
    # outward orientation: (B-A)x(C-B) points away from the centre
    a, b, c = xyz[tri[:, 0]], xyz[tri[:, 1]], xyz[tri[:, 2]]
    n = np.cross(b - a, c - b)
    assert np.all(np.einsum("ij,ij->i", n, (a + b + c) / 3) > 0)


def test_ellipsoid_and_plate():
    xyz, tri = hip.icosphere(2, 1.0, 0.5, 0.3)
    x2, t2 = meshgen.ellipsoid(2, 1.0, 0.5, 0.3)
    assert np.array_equal(tri, t2) and np.abs(xyz - x2).max() < 1e-15
    xyz, tri = hip.plate(6, 1.0)
    x2, t2 = meshgen.plate(6, 1.0)
    assert np.array_equal(tri, t2) and np.abs(xyz - x2).max() < 1e-15


@pytest.mark.parametrize("mesh", ["ico3", "plate7", "shuffled", "flipped"])
def test_rwg_builder_matches_reference_scan(mesh):
    rng = np.random.default_rng(0)
    if mesh == "ico3":
        _, tri = meshgen.icosphere(3, 1.0)
    elif mesh == "plate7":
        _, tri = meshgen.plate(7)
    elif mesh == "shuffled":
        _, tri = meshgen.icosphere(2, 1.0)
        tri = tri[rng.permutation(tri.shape[0])]
    else:  # a few triangles with the opposite orientation -> some edges are not shared consistently
        _, tri = meshgen.icosphere(2, 1.0)
        tri = tri.copy()
        tri[::17] = tri[::17, ::-1]
    r1, b1 = hip.rwg_build(tri)
    r2, b2 = O.build_rwg(tri)
    assert b1 == b2 and np.array_equal(r1, r2)


# ---------------------------------------------------------------- config
CONFIG_TEXT = """###########################################################
# test config: comments, duplicates, unknown keys
###########################################################
### [GLOBAL]
    EMSolver: EFIE
    result_type:   BistaticRCS
    mesh_path:   sphere.nas   # trailing comment
    frequency:   3e8
    polarization:   2
    Sigma: (0.0025,0)
### later duplicate wins
    polarization: Theta
    my_new_key: 42
"""


def _cfg(key, text=CONFIG_TEXT):
    L = ctypes.CDLL(CORE)
    buf = ctypes.create_string_buffer(256)
    rc = L.emfa_core_config_get(text.encode(), 1, key.encode(), buf, 256)
    return rc, buf.value.decode()


def test_config_semantics():
    assert _cfg("EMSolver:") == (0, "EFIE")
    assert _cfg("mesh_path:") == (0, "sphere.nas")
    assert _cfg("polarization:") == (0, "Theta")
    assert _cfg("Sigma:") == (0, "(0.0025,0)")
    assert _cfg("my_new_key:") == (0, "42")
    assert _cfg("EMSolver")[0] == 1  # keys keep their colon
    assert _cfg("frequency:") == (0, "3e8")


# ---------------------------------------------------------------- NAS reader
def write_nas(path, xyz, tri):
    """NASTRAN in the dialect EMComponent/Mesh.cpp:261-309 reads."""
    with open(path, "w") as f:
        for i in range(9):
            f.write(f"$ header line {i}\n")
        for i, p in enumerate(xyz):
            f.write(f"GRID*   {i + 1:8d}  {p[0]: .9e}  {p[1]: .9e}\n")
            f.write(f"*       {i + 1:8d}  {p[2]: .9e}\n")
        for i, t in enumerate(tri):
            f.write(f"CTRIA3 {i + 1} 1 {t[0] + 1} {t[1] + 1} {t[2] + 1}\n")
        f.write("ENDDATA\n")


def _read_mesh(path):
    L = ctypes.CDLL(CORE)
    nv, nt = ctypes.c_int(), ctypes.c_int()
    rc = L.emfa_core_read_mesh(path.encode(), None, 0, None, 0, ctypes.byref(nv), ctypes.byref(nt))
    if rc != 0:
        return rc, None, None
    xyz = np.zeros((nv.value, 3))
    tri = np.zeros((nt.value, 3), dtype=np.int32)
    rc = L.emfa_core_read_mesh(path.encode(), xyz.ctypes.data_as(ctypes.c_void_p), nv.value,
                               tri.ctypes.data_as(ctypes.c_void_p), nt.value, ctypes.byref(nv), ctypes.byref(nt))
    return rc, xyz, tri


def test_nas_round_trip(tmp_path):
    xyz, tri = meshgen.icosphere(2, 0.5)
    p = str(tmp_path / "s.nas")
    write_nas(p, xyz, tri)
    rc, x2, t2 = _read_mesh(p)
    assert rc == 0 and np.array_equal(t2, tri)
    # coordinates rounded to float with 6 significant digits, as the reference's temp.dat round trip
    expect = np.array([[float(np.float32(float("%g" % np.float32(v)))) for v in row] for row in xyz])
    assert np.array_equal(x2, expect)


def test_nas_errors(tmp_path):
    p = str(tmp_path / "bad.nas")
    with open(p, "w") as f:
        f.write("\n" * 9 + "GRID*   2  0.0  0.0\n*  2  0.0\nENDDATA\n")  # ids must start at 1
    assert _read_mesh(p)[0] == 3
    with open(p, "w") as f:
        f.write("\n" * 9 + "XYZ nonsense\n")  # neither GRID*, CTRIA3 nor E: rejected, not looped on
    assert _read_mesh(p)[0] == 3
    assert _read_mesh(str(tmp_path / "missing.nas"))[0] == 3
    assert _read_mesh(str(tmp_path / "m.stl"))[0] == 3


# ---------------------------------------------------------------- launcher
def test_launcher_config_errors(tmp_path):
    cfg = tmp_path / "CONFIG.yml"
    cfg.write_text("EMSolver: NOPE\n")
    r = subprocess.run([EXE], input=str(cfg) + "\n", stdout=subprocess.PIPE, text=True, timeout=60)
    assert r.returncode == 2 and "undefined EM solver type (NOPE)" in r.stdout
    r = subprocess.run([EXE], input=str(tmp_path / "none.yml") + "\n", stdout=subprocess.PIPE, text=True,
                       timeout=60)
    assert r.returncode == 2 and "fail opening config file" in r.stdout


def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


@pytest.mark.skipif(_has_gpu(), reason="checks the behaviour without a device")
def test_server_fails_loudly_without_device():
    from emfa_amd import serve
    with pytest.raises(hip.EmfaError):
        serve.Server(workers=2)


def test_workload_generators_are_deterministic_and_in_range():
    from emfa_amd import serve, workload
    a = workload.mixed_stream(200)
    b = workload.mixed_stream(200)
    assert all(x["subdiv"] == y["subdiv"] and x["freq"] == y["freq"] and np.array_equal(x["angles"], y["angles"])
               for x, y in zip(a, b))
    ns = [workload.unknowns(r) for r in a]
    assert min(ns) >= 500 and max(ns) <= 15000
    assert all(r["formulation"] == hip.EFIE for r in a if r["kind"] == serve.PLATE)
    assert all(1 <= len(r["angles"]) <= 64 for r in a)
    keys = [(r["kind"], r["subdiv"], r["axes"], r["freq"], r["formulation"]) for r in a]
    repeats = len(keys) - len(set(keys))
    assert 0.2 * len(a) <= repeats <= 0.4 * len(a)
    c = workload.batch_problems()
    assert len(c) == 256 and {r["subdiv"] for r in c} == {2, 3, 4}
    assert all(0.3 <= x <= 1.2 for r in c for x in r["axes"])
    assert all(100e6 <= r["freq"] <= 600e6 and r["formulation"] == hip.EFIE for r in c)


# ---------------------------------------------------------------- edge cases
def test_geometry_argument_errors_and_size_queries():
    for bad in [(-1, 1.0), (10, 1.0), (2, 0.0)]:
        with pytest.raises(hip.EmfaError) as e:
            hip.icosphere(*bad)
        assert e.value.code == hip.EMFA_E_INVALID
    nv, nt = ctypes.c_int(), ctypes.c_int()
    assert hip.lib().emfa_geom_icosphere(3, ctypes.c_double(1), ctypes.c_double(1), ctypes.c_double(1), None, 0,
                                         None, 0, ctypes.byref(nv), ctypes.byref(nt)) == hip.EMFA_OK
    assert (nv.value, nt.value) == (642, 1280)
    xyz = np.zeros((10, 3))
    tri = np.zeros((10, 3), dtype=np.int32)
    rc = hip.lib().emfa_geom_icosphere(3, ctypes.c_double(1), ctypes.c_double(1), ctypes.c_double(1),
                                       xyz.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 10,
                                       tri.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), 10, ctypes.byref(nv),
                                       ctypes.byref(nt))
    assert rc == hip.EMFA_E_INVALID and b"too small" in hip.lib().emfa_last_error()


def test_rwg_builder_degenerate_and_ragged_meshes():
    # one triangle: no interior edge, three boundary edges
    rwg, nb = hip.rwg_build(np.array([[0, 1, 2]], dtype=np.int32))
    assert rwg.shape[0] == 0 and nb == 3
    # two triangles sharing an edge: one RWG, plus/minus triangles in order
    rwg, nb = hip.rwg_build(np.array([[0, 1, 2], [1, 0, 3]], dtype=np.int32))
    assert rwg.tolist() == [[0, 1, 2, 3, 0, 1]] and nb == 4
    # open plates: 3c^2 - 2c unknowns, 4c boundary edges
    for c in (1, 2, 7):
        xyz, tri = hip.plate(c, 1.0)
        rwg, nb = hip.rwg_build(tri)
        assert rwg.shape[0] == 3 * c * c - 2 * c and nb == 4 * c
    # a strip with a triangle removed (ragged boundary) keeps the reference order
    xyz, tri = hip.plate(3, 1.0)
    keep = np.delete(tri, 7, axis=0)
    rwg, nb = hip.rwg_build(keep)
    ref = O.build_rwg(keep)[0]
    assert np.array_equal(rwg, ref)
    with pytest.raises(hip.EmfaError):
        hip.rwg_build(np.zeros((0, 3), dtype=np.int32))


def test_mlfma_restatement_rejects_single_level():
    from oracle import mlfma as OM
    xyz, tri = hip.icosphere(1, 0.1)
    rwg, _ = hip.rwg_build(tri)
    with pytest.raises(ValueError):
        OM.Tree(xyz, tri, rwg, 1e8, 0.5)  # the whole object fits one box of the threshold
