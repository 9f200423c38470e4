"""Writes the golden vectors of tests/golden/ref_*.npz by running the
reference's own EFIE/CFIE code (oracle/_ref/emfa_ref, built by
`make -C oracle -f ref.mk` from the reference sources) on small synthetic
meshes.  Each fixture holds the inputs (mesh as the reference read it, RWG
list, frequency, incidence) and the reference's outputs (complex<float> Z
rows, V, the LU solution I, the RCS cut, and the per-triangle current
magnitudes of CFIE::calculateSurfaceCurrent computed from that I; for
CFIE and MFIE also the reference's BiCGSTAB solutions of its own Z, V).

The ref_contour_*.npz fixtures hold the reference's half-RWG lists of the
IP-SIE solver with their contour lines and the surface part of every
half-RWG pair's near entry.

  python tests/golden/make_ref_golden.py [contour]
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


def rpt_text(xyz, tri):
    """.rpt as EMComponent/Mesh.cpp:178-218 reads it: two header lines,
    triangles, a separator and a header line, then nodes (id x y z i j)."""
    out = ["RPT mesh", "TRIANGLES"]
    out += [f"{i + 1} {t[0] + 1} {t[1] + 1} {t[2] + 1}" for i, t in enumerate(tri)]
    out += ["END TRIANGLES", "NODES"]
    out += [f"{i + 1} {p[0]:.9e} {p[1]:.9e} {p[2]:.9e} 0 0" for i, p in enumerate(xyz)]
    return "\n".join(out) + "\n"


def stl_text(xyz, tri):
    """ASCII STL as EMComponent/Mesh.cpp:220-259 reads it."""
    out = ["solid emfa"]
    for t in tri:
        a, b, c = xyz[t[0]], xyz[t[1]], xyz[t[2]]
        n = np.cross(b - a, c - b)
        n = n / np.linalg.norm(n)
        out.append(f"  facet normal {n[0]:.9e} {n[1]:.9e} {n[2]:.9e}")
        out.append("    outer loop")
        for p in (a, b, c):
            out.append(f"      vertex {p[0]:.9e} {p[1]:.9e} {p[2]:.9e}")
        out.append("    endloop")
        out.append("  endfacet")
    out.append("endsolid emfa")
    return "\n".join(out) + "\n"


# name: (format, mesh builder)
MESH_CASES = {
    "rpt_sphere_s1": ("rpt", lambda: meshgen.icosphere(1, 0.731)),
    "rpt_plate3": ("rpt", lambda: meshgen.plate(3, 1.3)),
    "stl_sphere_s1": ("stl", lambda: meshgen.icosphere(1, 0.731)),
    "stl_ellipsoid_s1": ("stl", lambda: meshgen.ellipsoid(1, 1.1, 0.7, 0.45)),
}


def read_mesh_dump(path):
    with open(path, "rb") as f:
        nv, nt = np.frombuffer(f.read(8), dtype=np.int32)
        xyz = np.frombuffer(f.read(12 * nv), dtype=np.float32).reshape(nv, 3)
        rec = np.frombuffer(f.read(60 * nt), dtype=np.uint8).reshape(nt, 60)
    ids = rec[:, :12].copy().view(np.int32).reshape(nt, 3)
    fl = rec[:, 12:].copy().view(np.float32).reshape(nt, 12)
    return xyz, ids, fl[:, :9].reshape(nt, 3, 3), fl[:, 9:]


def make_mesh_fixtures():
    for name, (fmt, build) in MESH_CASES.items():
        xyz, tri = build()
        text = rpt_text(xyz, tri) if fmt == "rpt" else stl_text(xyz, tri)
        with tempfile.TemporaryDirectory() as d:
            path = os.path.join(d, "mesh." + fmt)
            with open(path, "w") as f:
                f.write(text)
            out = os.path.join(d, "out.bin")
            subprocess.run([EXE, "mesh", path, out], cwd=d, check=True, stdout=subprocess.DEVNULL)
            rx, ids, corners, normals = read_mesh_dump(out)
        np.savez_compressed(os.path.join(HERE, f"ref_mesh_{name}.npz"), format=fmt, text=np.array(text),
                            xyz=rx, ids=ids, corners=corners, normals=normals)
        print("mesh", name, "vertices", rx.shape[0], "triangles", ids.shape[0])


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


# name: (mesh builder, freq, alpha, rows of the half-RWG surface entries kept)
CONTOUR_CASES = {
    "sphere_s1": (lambda: meshgen.icosphere(1, 0.5), 3e8, 0.5, 64),
    "cube_232322": (lambda: meshgen.cube_faces([2, 3, 2, 3, 2, 2], 1.0), 3e8, 0.5, None),
    "plate_pair_34": (lambda: meshgen.plate_pair(3, 4, 1.0), 4e8, 0.3, None),
}


def read_contour(path):
    with open(path, "rb") as f:
        nv, nt, n, nl = np.frombuffer(f.read(16), dtype=np.int32)
        xyz = np.frombuffer(f.read(12 * nv), dtype=np.float32).reshape(nv, 3)
        tri = np.frombuffer(f.read(12 * nt), dtype=np.int32).reshape(nt, 3)
        rec = np.frombuffer(f.read(28 * n), dtype=np.uint8).reshape(n, 28)
        hrwg = rec[:, :24].copy().view(np.int32).reshape(n, 6)
        length = rec[:, 24:].copy().view(np.float32).reshape(n)
        ptr, ltri, seg = [0], [], []
        for _ in range(nl):
            (c,) = np.frombuffer(f.read(4), dtype=np.int32)
            for _ in range(c):
                r = f.read(28)
                ltri.append(int(np.frombuffer(r[:4], dtype=np.int32)[0]))
                seg.append(np.frombuffer(r[4:], dtype=np.float32))
            ptr.append(len(ltri))
        z = np.frombuffer(f.read(8 * n * n), dtype=np.float32).reshape(n, n, 2)
    seg = np.array(seg, dtype=np.float32).reshape(-1, 6)
    return xyz, tri, hrwg, length, np.array(ptr, dtype=np.int32), np.array(ltri, dtype=np.int32), seg, \
        z[..., 0].astype(np.float64) + 1j * z[..., 1].astype(np.float64)


def make_contour_fixtures():
    """Half-RWG lists of ContourEdge::buildEdge(mesh, HALF_RWG) with their
    contour lines, and the surface part of every half-RWG pair's near entry
    (EMSolver/FastWT.cpp:753-762 around Miscellaneous.cpp's kernels)."""
    for name, (build, freq, alpha, rows) in CONTOUR_CASES.items():
        xyz, tri = build()
        with tempfile.TemporaryDirectory() as d:
            nas = os.path.join(d, "mesh.nas")
            write_nas(nas, xyz, tri)
            out = os.path.join(d, "out.bin")
            subprocess.run([EXE, "contour", nas, repr(freq), repr(alpha), out], cwd=d, check=True,
                           stdout=subprocess.DEVNULL)
            rx, rt, hrwg, length, ptr, ltri, seg, z = read_contour(out)
        keep = z if rows is None else z[:rows]
        np.savez_compressed(os.path.join(HERE, f"ref_contour_{name}.npz"), freq=freq, alpha=alpha, xyz=rx, tri=rt,
                            hrwg=hrwg, length=length, line_ptr=ptr, line_tri=ltri, line_seg=seg,
                            zfs=keep.astype(np.complex64))
        print("contour", name, "half-RWGs", hrwg.shape[0], "nonconformal", len(ptr) - 1, "lines", len(ltri))


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
    if sys.argv[1:] == ["contour"]:
        make_contour_fixtures()
    else:
        main()
        make_mesh_fixtures()
        make_contour_fixtures()
