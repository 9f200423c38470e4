"""The drop-in path end to end on the device: CONFIG.yml + NASTRAN mesh ->
emfa launcher (solver::loadSolver -> init/solve/output/clear) -> the
reference's result files, compared with the FP64 oracle on the same mesh."""
import os
import subprocess
import sys

import numpy as np
import pytest

import meshgen
from conftest import ROOT
from oracle import oracle as O
from test_cpu_host import _read_mesh, write_nas

pytestmark = pytest.mark.gpu
EXE = os.path.join(ROOT, "emfa_amd", "bin", "emfa")

BASE = """    EMSolver: {solver}
    result_type: {rtype}
    mesh_path: {mesh}
    frequency: {freq}
    pole: 0
    frequency_from: 6e8
    frequency_to: 6e8
    frequency_delta: 6e8
    frequency_inc: 6e8
    theta: 0
    phi: 0
    theta_from: 0
    theta_to: {tto}
    theta_delta: {tdelta}
    phi_from: 0
    phi_to: 90
    phi_delta: 90
    E_Plane: true
    H_Plane: true
    HH_Plane: true
    Origin_point: (0,0,1)
    End_point: (3,3,1)
    Sampling_x: 1
    Sampling_y: 1
    Sampling_z: 1
    alpha: 0.5
    Sigma: (0,0)
    Dx: 0.6
    Dy: 0.6
    Array_x: 20
    Array_y: 20
"""


def _run(tmp_path, solver, rtype, extra="", freq="3e8", radius=0.5, subdiv=2, tto=360, tdelta=5, geometry=None,
         name="sphere"):
    xyz, tri = meshgen.icosphere(subdiv, radius) if geometry is None else geometry
    mesh = str(tmp_path / f"{name}.nas")
    write_nas(mesh, xyz, tri)
    cfg = tmp_path / "CONFIG.yml"
    cfg.write_text("# drop-in test\n" + BASE.format(solver=solver, rtype=rtype, mesh=mesh, freq=freq, tto=tto,
                                                    tdelta=tdelta) + extra)
    print(f"launcher {solver} {rtype} {extra.strip()}", flush=True)
    r = subprocess.run([EXE], input=str(cfg) + "\n", stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout[-3000:]
    assert "Exit State: OK" in r.stdout
    rc, xr, tr = _read_mesh(mesh)
    assert rc == 0
    return xr, tr


def _read_rcs(path):
    lines = open(path).read().splitlines()
    assert lines[0].split() == ["Theta", "RCS(dBm^2)"]
    a = np.array([[float(v) for v in ln.split()] for ln in lines[1:]])
    return a[:, 0], a[:, 1]


def _dirs(th, ph):
    th, ph = np.radians(th), np.radians(ph)
    return np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], 1)


@pytest.mark.parametrize("precision", ["FP64", "FP32"])
def test_efie_bistatic_files_match_oracle(tmp_path, precision):
    xyz, tri = _run(tmp_path, "EFIE", "BistaticRCS", extra=f"    precision: {precision}\n")
    out = tmp_path / "sphere_EFIE"
    for f in ["bistatic_rcs_phi_0.dat", "bistatic_rcs_phi_90.dat", "bistatic_rcs_theta_90.dat", "current.sc",
              "sca_runtime.log"]:
        assert (out / f).exists(), f
    rwg, _ = O.build_rwg(tri)
    k = O.wavenumber(3e8)
    Z = O.fill_efie(xyz, tri, rwg, k, 3)
    X = np.linalg.solve(Z, O.rhs(xyz, tri, rwg, 0, k, 1.0, O.inc_table([(0, 0, 0)])))
    tol = 0.05 if precision == "FP64" else 0.1
    for ph in (0, 90):
        th, val = _read_rcs(out / f"bistatic_rcs_phi_{ph}.dat")
        assert np.array_equal(th, np.arange(0, 361, 5.0))
        ref = O.rcs(xyz, tri, rwg, k, X, _dirs(th, ph))
        assert np.abs(val - ref).max() <= tol + 1e-4  # files carry 4 decimals
    ph, val = _read_rcs(out / "bistatic_rcs_theta_90.dat")
    assert len(ph) == 360
    ref = O.rcs(xyz, tri, rwg, k, X, _dirs(np.full(360, 90.0), ph))
    assert np.abs(val - ref).max() <= tol + 1e-4
    # current.sc: header + one line per triangle with 13 numbers
    lines = open(out / "current.sc").read().splitlines()
    assert lines[0] == f"triangles: {tri.shape[0]}" and len(lines) == tri.shape[0] + 2
    mags, _ = O.current(xyz, tri.shape[0], rwg, X)
    vals = np.array([[float(v) for v in ln.split()[1:]] for ln in lines[2:]])
    assert np.allclose(vals[:, 9:13], mags, rtol=2e-4 if precision == "FP64" else 1e-3, atol=1e-9)


def test_cfie_monostatic_multi_rhs(tmp_path):
    xyz, tri = _run(tmp_path, "CFIE", "MonostaticRCS", freq="6e8", radius=1.0, tto=180, tdelta=30,
                    extra="    precision: FP32\n    quadrature_order: 3\n    linear_solver: LU\n")
    out = tmp_path / "sphere_CFIE_0.5"
    rwg, _ = O.build_rwg(tri)
    k = O.wavenumber(6e8)
    Z = O.fill_cfie(xyz, tri, rwg, k, 0.5, 3, 0)
    for ph in (0, 90):
        th, val = _read_rcs(out / f"monostatic_rcs_phi_{ph}.dat")
        B = O.rhs(xyz, tri, rwg, 1, k, 0.5, O.inc_table([(t, ph, 0) for t in th]))
        X = np.linalg.solve(Z, B)
        ref = O.rcs(xyz, tri, rwg, k, X, _dirs(th, ph), np.arange(len(th)))
        assert np.abs(val - ref).max() <= 0.1 + 1e-4


def test_cfie_default_bicgstab_as_the_reference(tmp_path):
    """EMSolver: CFIE without linear_solver: solves with BiCGSTAB at the
    reference's settings (CFIE.cpp:76: eps 0.01, at most 500 steps, x0 = 0),
    on the device's FP64 matrix.  The oracle's restatement of the same
    iml::BICGSTAB overload on the oracle's matrix takes the same steps and
    gives the same RCS to the files' 4 decimals; the log carries the
    reference's "Iterative number" line."""
    xyz, tri = _run(tmp_path, "CFIE", "BistaticRCS", freq="6e8", radius=1.0, tto=180, tdelta=10)
    out = tmp_path / "sphere_CFIE_0.5"
    log = open(out / "sca_runtime.log").read()
    assert "linear solver: BICGSTAB" in log and "Iterative number" in log
    rwg, _ = O.build_rwg(tri)
    k = O.wavenumber(6e8)
    Z = O.fill_cfie(xyz, tri, rwg, k, 0.5, 3, 0)
    b = O.rhs(xyz, tri, rwg, 1, k, 0.5, O.inc_table([(0, 0, 0)]))[:, 0]
    x, steps, ok = O.bicgstab(Z, b, 0.01, 500)
    assert ok and f"Iterative number: {steps}," in log
    th, val = _read_rcs(out / "bistatic_rcs_phi_0.dat")
    ref = O.rcs(xyz, tri, rwg, k, x, _dirs(th, 0))
    assert np.abs(val - ref).max() <= 1e-3


def test_mfie_and_gmres(tmp_path):
    xyz, tri = _run(tmp_path, "MFIE", "BistaticRCS", freq="3e8", radius=0.5,
                    extra="    linear_solver: GMRES\n    gmres_tolerance: 1e-10\n")
    out = tmp_path / "sphere_MFIE"
    rwg, _ = O.build_rwg(tri)
    k = O.wavenumber(3e8)
    Z = O.fill_cfie(xyz, tri, rwg, k, 0.0, 3, 0)
    X = np.linalg.solve(Z, O.rhs(xyz, tri, rwg, 2, k, 0.0, O.inc_table([(0, 0, 0)])))
    th, val = _read_rcs(out / "bistatic_rcs_phi_0.dat")
    ref = O.rcs(xyz, tri, rwg, k, X, _dirs(th, 0))
    assert np.abs(val - ref).max() <= 0.05 + 1e-4


def test_write_matrix(tmp_path):
    xyz, tri = _run(tmp_path, "EFIE", "BistaticRCS", extra="    write_matrix: true\n", subdiv=1, tto=10)
    out = tmp_path / "sphere_EFIE"
    lines = open(out / "matrix_Z.txt").read().splitlines()
    rwg, _ = O.build_rwg(tri)
    n = rwg.shape[0]
    assert lines[0] == "ARMA_MAT_TXT_FC008" and lines[1] == f"{n} {n}" and len(lines) == n + 2
    row0 = [complex(*map(float, t.strip("()").split(","))) for t in lines[2].split()]
    Z = O.fill_efie(xyz, tri, rwg, O.wavenumber(3e8), 3)
    assert np.allclose(row0, Z[0], rtol=1e-5, atol=1e-5 * np.abs(Z).max())
    assert (out / "matrix_V.txt").exists() and (out / "martix_I.txt").exists()


def test_mlfma_solver_files_near_dense_cfie(tmp_path):
    # EMSolver: MLFMA on a 4-times subdivided sphere (7680 unknowns, 4 levels),
    # against the dense CFIE solution of the oracle on the same mesh
    xyz, tri = _run(tmp_path, "MLFMA", "BistaticRCS", freq="8e8", radius=0.5, subdiv=4, tto=180, tdelta=10,
                    extra="    iteration_threshold: 1e-4\n    max_iteration_number: 600\n"
                          "    mlfma_box_threshold: 0.25\n    preconditioning: false\n")
    out = tmp_path / "sphere_MLFMA_CFIE_0.5"
    log = open(out / "sca_runtime.log").read()
    assert "levels: 4" in log and "Iterative number" in log and "cannot converge" not in log
    rwg, _ = O.build_rwg(tri)
    k = O.wavenumber(8e8)
    Z = O.fill_cfie(xyz, tri, rwg, k, 0.5, 3, 0)
    X = np.linalg.solve(Z, O.rhs(xyz, tri, rwg, 1, k, 0.5, O.inc_table([(0, 0, 0)])))
    th, val = _read_rcs(out / "bistatic_rcs_phi_0.dat")
    ref = O.rcs(xyz, tri, rwg, k, X, _dirs(th, 0))
    # the multipole operator's own accuracy on this mesh (a few per cent):
    # 1 dB where the pattern is within 15 dB of its peak, 5 % in linear power
    lit = ref > ref.max() - 15
    assert np.abs(val - ref)[lit].max() <= 1.0
    p, q = 10 ** (val / 10), 10 ** (ref / 10)
    assert np.linalg.norm(p - q) / np.linalg.norm(q) < 5e-2


def test_mlfma_preconditioned_solver_files(tmp_path):
    # preconditioning: true with the shipped thresholds: same RCS as the
    # unpreconditioned operator (dense CFIE check above), fewer iterations
    xyz, tri = _run(tmp_path, "MLFMA", "BistaticRCS", freq="8e8", radius=0.5, subdiv=4, tto=180, tdelta=10,
                    extra="    iteration_threshold: 1e-4\n    max_iteration_number: 600\n"
                          "    mlfma_box_threshold: 0.25\n    preconditioning: true\n"
                          "    row_threshold: 0.9\n    col_threshold: 0.87\n")
    out = tmp_path / "sphere_MLFMA_CFIE_0.5"
    log = open(out / "sca_runtime.log").read()
    assert "Number of precondition" in log and "cannot converge" not in log
    rwg, _ = O.build_rwg(tri)
    k = O.wavenumber(8e8)
    Z = O.fill_cfie(xyz, tri, rwg, k, 0.5, 3, 0)
    X = np.linalg.solve(Z, O.rhs(xyz, tri, rwg, 1, k, 0.5, O.inc_table([(0, 0, 0)])))
    th, val = _read_rcs(out / "bistatic_rcs_phi_0.dat")
    ref = O.rcs(xyz, tri, rwg, k, X, _dirs(th, 0))
    lit = ref > ref.max() - 15
    assert np.abs(val - ref)[lit].max() <= 1.0


def test_fmm_solver_files_near_dense_cfie(tmp_path):
    # EMSolver: FMM (single level, fmm_box_length: 0.4 wavelengths) on the
    # 1920-unknown sphere at 600 MHz, BiCGSTAB as FMM.cpp:222-268
    xyz, tri = _run(tmp_path, "FMM", "BistaticRCS", freq="6e8", radius=0.5, subdiv=3, tto=180, tdelta=10,
                    extra="    iteration_threshold: 1e-4\n    max_iteration_number: 600\n"
                          "    fmm_box_length: 0.4\n")
    out = tmp_path / "sphere_FMM_CFIE_0.5"
    log = open(out / "sca_runtime.log").read()
    assert "Valid boxes" in log and "Iterative number" in log and "cannot converge" not in log
    boxes = open(out / "ffm_box_log.txt").read().splitlines()
    assert sum(1 for l in boxes if l.startswith("    -> Box ")) > 1
    rwg, _ = O.build_rwg(tri)
    k = O.wavenumber(6e8)
    Z = O.fill_cfie(xyz, tri, rwg, k, 0.5, 3, 0)
    X = np.linalg.solve(Z, O.rhs(xyz, tri, rwg, 1, k, 0.5, O.inc_table([(0, 0, 0)])))
    th, val = _read_rcs(out / "bistatic_rcs_phi_0.dat")
    ref = O.rcs(xyz, tri, rwg, k, X, _dirs(th, 0))
    lit = ref > ref.max() - 15
    assert np.abs(val - ref)[lit].max() <= 0.5


def test_launcher_beside_a_process_that_launched_cooperatively(tmp_path):
    # This process factors a matrix on a plain context first (cooperative
    # panel launches) and keeps the context open; the launcher process that
    # follows may then have its cooperative launches refused and must fall
    # back to normal panel launches, with the same results.
    from emfa_amd import hip
    rng = np.random.default_rng(11)
    n = 1000
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    ctx = hip.Context(0)
    held = []
    try:
        Z = hip.Mat.from_numpy(ctx, A, hip.F64)
        held.append(Z)
        held.append(hip.LU(ctx, Z, hip.F64))
        ctx.sync()
        xyz, tri = _run(tmp_path, "EFIE", "BistaticRCS", extra="    precision: FP64\n", tto=60, tdelta=10)
    finally:
        for x in reversed(held):
            x.close()
        ctx.close()
    rwg, _ = O.build_rwg(tri)
    k = O.wavenumber(3e8)
    Z = O.fill_efie(xyz, tri, rwg, k, 3)
    X = np.linalg.solve(Z, O.rhs(xyz, tri, rwg, 0, k, 1.0, O.inc_table([(0, 0, 0)])))
    th, val = _read_rcs(tmp_path / "sphere_EFIE" / "bistatic_rcs_phi_0.dat")
    assert np.array_equal(th, np.arange(0, 61, 10.0))
    ref = O.rcs(xyz, tri, rwg, k, X, _dirs(th, 0))
    assert np.abs(val - ref).max() <= 0.05 + 1e-4


def test_fastwt_solver_files(tmp_path):
    """EMSolver: FastWT on a cube whose faces meet nonconformally (4 and 5
    cells per edge): the files of FastWT::output, the RCS against the
    restated IP-SIE operator's own answer (GMRES on oracle/fastwt.py's
    operator) and current.sc against the restated current of that answer."""
    from emfa_amd import hip
    from oracle import fastwt as W
    from scipy.sparse.linalg import LinearOperator, gmres
    geo = meshgen.cube_faces([4, 5, 4, 5, 4, 5], 0.6)
    xyz, tri = _run(tmp_path, "FastWT", "BistaticRCS", freq="3e8", tto=180, tdelta=10, geometry=geo, name="cube",
                    extra="    iteration_threshold: 1e-6\n    max_iteration_number: 2000\n"
                          "    mlfma_box_threshold: 0.2\n    stabilization_factor: 0.1\n    average_size: 0.07503\n")
    out = tmp_path / "cube_FastWT_0.5"
    for f in ["bistatic_rcs_phi_0.dat", "bistatic_rcs_phi_90.dat", "bistatic_rcs_theta_90.dat", "current.sc",
              "sca_parallel_runtime.log", "level_detail_info.txt"]:
        assert (out / f).exists(), f
    log = open(out / "sca_parallel_runtime.log").read()
    assert "nonconformal" in log and "Iterative number" in log and "cannot converge" not in log
    c = hip.contour_build(xyz, tri)
    C = W.Contour(c.hrwg, c.length, c.line_ptr, c.line_tri, c.line_seg)
    assert (C.hrwg[:, 5] == 1).sum() > 50
    f, k = 3e8, O.wavenumber(3e8)
    T = W.tree(xyz, tri, C, f, 0.2)
    Zn = W.near_entries(xyz, tri, C, T, k, 0.5, W.beta(0.1, 0.07503))
    rad, rec = W.patterns(T, xyz, tri, C, 0.5)
    A = LinearOperator((C.n, C.n), matvec=lambda x: W.matvec(T, Zn, rad, rec, x), dtype=complex)
    B = W.rhs(xyz, tri, C, k, 0.5, O.inc_table([(0, 0, 0)]))[:, 0]
    x, info = gmres(A, B, rtol=1e-8, restart=300, maxiter=5)
    assert info == 0
    th, val = _read_rcs(out / "bistatic_rcs_phi_0.dat")
    ref = W.rcs(xyz, tri, C, k, x, _dirs(th, 0))
    lit = ref > ref.max() - 15
    assert np.abs(val - ref)[lit].max() <= 0.05
    lines = open(out / "current.sc").read().splitlines()
    assert lines[0] == f"triangles: {tri.shape[0]}" and len(lines) == tri.shape[0] + 2
    mags, _ = W.current(xyz, tri, C, x)
    vals = np.array([[float(v) for v in ln.split()[1:]] for ln in lines[2:]])
    assert np.allclose(vals[:, 9:13], mags, rtol=2e-3, atol=1e-6 * mags.max())
