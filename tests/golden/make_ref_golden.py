"""Writes the golden vectors of tests/golden/ref_*.npz by running the
reference's own EFIE/CFIE code (oracle/_ref/emfa_ref, built by
`make -C oracle -f ref.mk` from the reference sources) on small synthetic
meshes.  Each fixture holds the inputs (mesh as the reference read it, RWG
list, frequency, incidence) and the reference's outputs (complex<float> Z
rows, V, the LU solution I, the RCS cut, and the per-triangle current
magnitudes of CFIE::calculateSurfaceCurrent computed from that I; for
CFIE and MFIE also the reference's BiCGSTAB solutions of its own Z, V).

  python tests/golden/make_ref_golden.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(HERE))
import meshgen  # noqa: E402
from test_cpu_host import write_nas  # noqa: E402

EXE = os.path.join(ROOT, "oracle", "_ref", "emfa_ref")

# name: (kind, mesh builder, freq, alpha, sigma, (theta, phi, pole), rows of Z kept)
CASES = {
    "efie_sphere_s1": ("efie", lambda: meshgen.icosphere(1, 0.5), 3e8, 1.0, 0j, (0, 0, 0), None),
    "efie_sphere_s2_sigma": ("efie", lambda: meshgen.icosphere(2, 0.5), 3e8, 1.0, 0.0025 + 0.001j, (20, 10, 30), 64),
    "efie_plate4": ("efie", lambda: meshgen.plate(4, 1.0), 3e8, 1.0, 0j, (30, 45, 0), None),
    "cfie_sphere_s2": ("cfie", lambda: meshgen.icosphere(2, 1.0), 6e8, 0.5, 0j, (0, 0, 0), 64),
    "cfie_ellipsoid_s1": ("cfie", lambda: meshgen.ellipsoid(1, 1.0, 0.6, 0.4), 4e8, 0.3, 0j, (40, 20, 90), None),
    "mfie_sphere_s1": ("mfie", lambda: meshgen.icosphere(1, 0.5), 3e8, 0.0, 0j, (0, 0, 0), None),
    "mfie_ellipsoid_s1": ("mfie", lambda: meshgen.ellipsoid(1, 0.8, 0.5, 0.6), 4.5e8, 0.0, 0j, (60, 30, 45), None),
}


def read_out(path):
    with open(path, "rb") as f:
        nv, nt, n, na = np.frombuffer(f.read(16), dtype=np.int32)
        xyz = np.frombuffer(f.read(12 * nv), dtype=np.float32).reshape(nv, 3)
        tri = np.frombuffer(f.read(12 * nt), dtype=np.int32).reshape(nt, 3)
        rwg = np.frombuffer(f.read(24 * n), dtype=np.int32).reshape(n, 6)
        Z = np.frombuffer(f.read(8 * n * n), dtype=np.float32).reshape(n, n, 2)
        V = np.frombuffer(f.read(8 * n), dtype=np.float32).reshape(n, 2)
        I = np.frombuffer(f.read(8 * n), dtype=np.float32).reshape(n, 2)
        th = np.frombuffer(f.read(4 * na), dtype=np.float32)
        rcs = np.frombuffer(f.read(4 * na), dtype=np.float32)
        (nc,) = np.frombuffer(f.read(4), dtype=np.int32)
        cur = np.frombuffer(f.read(52 * nc), dtype=np.float32).reshape(nc, 13)
        it_b, it_mf = np.frombuffer(f.read(8), dtype=np.int32)
        xb = xmf = None
        if it_b >= 0:
            xb = np.frombuffer(f.read(8 * n), dtype=np.float32).reshape(n, 2)
            xmf = np.frombuffer(f.read(8 * n), dtype=np.float32).reshape(n, 2)
    c = lambda a: a[..., 0].astype(np.float64) + 1j * a[..., 1].astype(np.float64)
    bicg = None if xb is None else (int(it_b), c(xb), int(it_mf), c(xmf))
    return xyz, tri, rwg, c(Z), c(V), c(I), th, rcs, cur, bicg


def main():
    if not os.path.exists(EXE):
        sys.exit("oracle/_ref/emfa_ref is not built (make -C oracle -f ref.mk)")
    for name, (kind, build, freq, alpha, sigma, inc, rows) in CASES.items():
        xyz, tri = build()
        with tempfile.TemporaryDirectory() as d:
            nas = os.path.join(d, "mesh.nas")
            write_nas(nas, xyz, tri)
            out = os.path.join(d, "out.bin")
            subprocess.run([EXE, kind, nas, repr(freq), repr(alpha), repr(sigma.real), repr(sigma.imag),
                            str(inc[0]), str(inc[1]), str(inc[2]), out], cwd=d, check=True,
                           stdout=subprocess.DEVNULL)
            rx, rt, rr, Z, V, I, th, rcs, cur, bicg = read_out(out)
        keep = Z if rows is None else Z[:rows]
        extra = {}
        if bicg is not None:
            # BiCGSTAB of the reference on its own Z and V (both overloads; the
            # matrix-free one never writes its step count, EMIterative/iml.cpp:88-154)
            extra = dict(bicgstab_steps=bicg[0], bicgstab_x=bicg[1].astype(np.complex64),
                         bicgstab_mf_x=bicg[3].astype(np.complex64))
        np.savez_compressed(os.path.join(HERE, f"ref_{name}.npz"), kind=kind, freq=freq, alpha=alpha,
                            sigma=sigma, inc=np.array(inc, dtype=np.float64), xyz=rx, tri=rt, rwg=rr,
                            Z=keep.astype(np.complex64), V=V.astype(np.complex64), I=I.astype(np.complex64),
                            theta=th, rcs=rcs, cur_vertices=cur[:, :9], cur_magn=cur[:, 9:], **extra)
        print(name, "N =", rr.shape[0], "Z rows kept:", keep.shape[0])


if __name__ == "__main__":
    main()
