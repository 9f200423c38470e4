"""Synthetic request streams and the two ways of answering them.

* `batch_problems` — many uneven EFIE problems: ellipsoids with semi-axes
  U[0.3, 1.2] m, icosphere subdivisions {2, 3, 4} (320..5120 triangles),
  100..600 MHz, 1..64 incidences each (theta spread over 0..180 deg).
* `mixed_stream` — spheres, ellipsoids and 1 x 1 m plates (2 triangles per
  cell) with 500..15000 RWG unknowns (plate sizes stratified over the range),
  EFIE/CFIE (plates EFIE only: CFIE needs a closed surface), 1..64
  incidences, 30 % of the requests repeat an earlier (mesh, frequency,
  formulation) with new incidences.

`run_plain` answers a list of requests one after another through the C ABI
with a fresh FP64 fill and LU each (the drop-in path, nothing cached);
`run_served` sends them through the solve server (worker contexts, HBM
admission, resident factors), either all at once or on Poisson arrivals.
"""
import threading
import time

import numpy as np

from . import hip
from . import serve

SEED = 20261202


def _angles(rng, nrhs, theta_max):
    th = np.linspace(0.0, theta_max, nrhs) if nrhs > 1 else np.array([rng.uniform(0.0, theta_max)])
    return np.stack([th, np.zeros(nrhs), np.zeros(nrhs)], 1)


def batch_problems(count=256, seed=SEED):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        axes = tuple(float(x) for x in rng.uniform(0.3, 1.2, 3))
        s = int(rng.integers(2, 5))
        f = float(rng.uniform(100e6, 600e6))
        nrhs = int(rng.integers(1, 65))
        out.append(dict(kind=serve.ICOSPHERE, subdiv=s, axes=axes, formulation=hip.EFIE, alpha=1.0, freq=f,
                        angles=_angles(rng, nrhs, 180.0)))
    return out


def _plate_cells(n_unknowns):
    """Cells per side of the square plate with about n_unknowns RWGs (3c^2 - 2c)."""
    return int(round((2.0 + np.sqrt(4.0 + 12.0 * n_unknowns)) / 6.0))


def mixed_stream(count, seed=SEED, repeat=0.3):
    """New plates take their size from four equal strata of 500..15000
    unknowns in turn (a size drawn uniformly inside the stratum), so that any
    prefix of the stream covers the whole range."""
    rng = np.random.default_rng(seed)
    out = []
    new_plates = 0
    for _ in range(count):
        nrhs = int(rng.integers(1, 65))
        if out and rng.random() < repeat:
            base = out[int(rng.integers(len(out)))]
            r = dict(base)
            r["angles"] = _angles(rng, nrhs, 80.0 if base["kind"] == serve.PLATE else 180.0)
            out.append(r)
            continue
        fam = int(rng.integers(3))
        f = float(rng.uniform(100e6, 600e6))
        if fam == 2:  # plate: 3c^2 - 2c unknowns, 500..15000
            lo = 500.0 + 3625.0 * (new_plates % 4)
            c = min(71, max(14, _plate_cells(rng.uniform(lo, lo + 3625.0))))
            new_plates += 1
            out.append(dict(kind=serve.PLATE, subdiv=c, axes=(1.0, 1.0, 1.0), formulation=hip.EFIE, alpha=1.0,
                            freq=f, angles=_angles(rng, nrhs, 80.0)))
        else:  # sphere / ellipsoid, subdivisions 3..4 (1920 / 7680 unknowns)
            s = int(rng.integers(3, 5))
            if fam == 0:
                a = float(rng.uniform(0.3, 1.2))
                axes = (a, a, a)
            else:
                axes = tuple(float(x) for x in rng.uniform(0.3, 1.2, 3))
            cfie = rng.random() < 0.5
            out.append(dict(kind=serve.ICOSPHERE, subdiv=s, axes=axes, formulation=hip.CFIE if cfie else hip.EFIE,
                            alpha=0.5 if cfie else 1.0, freq=f, angles=_angles(rng, nrhs, 180.0)))
    return out


def nonconformal_cube(res, size=1.0):
    """Closed cube surface of edge `size` centred at the origin, its six faces
    (+x, -x, +y, -y, +z, -z) meshed separately with res[f] x res[f] cells,
    each face with its own vertices, triangles counter-clockwise seen from
    outside: faces of unequal resolution meet nonconformally, equal ones with
    coincident vertices under distinct ids (the IP-SIE solver's workload)."""
    h = size / 2
    s = 2 * h
    faces = [((h, -h, -h), (0, s, 0), (0, 0, s)), ((-h, -h, -h), (0, 0, s), (0, s, 0)),
             ((-h, h, -h), (0, 0, s), (s, 0, 0)), ((-h, -h, -h), (s, 0, 0), (0, 0, s)),
             ((-h, -h, h), (s, 0, 0), (0, s, 0)), ((-h, -h, -h), (0, s, 0), (s, 0, 0))]
    xyz, tri, off = [], [], 0
    for r, (o, a, b) in zip(res, faces):
        o, a, b = (np.asarray(v, dtype=np.float64) for v in (o, a, b))
        i, j = np.meshgrid(np.arange(r + 1), np.arange(r + 1))
        pts = o + (i.ravel()[:, None] / r) * a + (j.ravel()[:, None] / r) * b
        jj, ii = np.meshgrid(np.arange(r), np.arange(r), indexing="ij")
        v00 = (jj * (r + 1) + ii).ravel()
        v10, v01 = v00 + 1, v00 + r + 1
        t = np.stack([np.stack([v00, v10, v01 + 1], 1), np.stack([v00, v01 + 1, v01], 1)], 1).reshape(-1, 3)
        xyz.append(np.round(pts, 9))
        tri.append(t + off)
        off += len(pts)
    return np.concatenate(xyz), np.concatenate(tri).astype(np.int32)


def unknowns(r):
    return serve.unknowns(r["kind"], r["subdiv"])


def _mesh(r):
    if r["kind"] == serve.PLATE:
        xyz, tri = hip.plate(r["subdiv"], r["axes"][0])
    else:
        xyz, tri = hip.icosphere(r["subdiv"], *r["axes"])
    rwg, _ = hip.rwg_build(tri)
    return xyz, tri, rwg


def solve_plain(ctx, r, quad=3):
    """One request on the plain FP64 path: mesh, FP64 fill, FP64 LU, solve,
    monostatic RCS.  Returns the RCS (dB) per incidence."""
    xyz, tri, rwg = _mesh(r)
    n = rwg.shape[0]
    k = hip.wavenumber(r["freq"])
    inc = hip.inc_table(r["angles"])
    dirs = -inc[:, 0:3]
    mesh = hip.Mesh(ctx, xyz, tri, rwg)
    Z = hip.Mat(ctx, n, n, hip.F64)
    B = hip.Mat(ctx, n, inc.shape[0], hip.F64)
    try:
        hip.fill(ctx, mesh, Z, r["formulation"], k, quad, alpha=r["alpha"])
        hip.rhs_fill(ctx, mesh, B, r["formulation"], k, r["alpha"], inc)
        lu = hip.LU(ctx, Z, hip.F64)
        try:
            lu.solve(B)
            return hip.rcs(ctx, mesh, k, B, dirs, np.arange(inc.shape[0], dtype=np.int32))
        finally:
            lu.close()
    finally:
        B.close()
        Z.close()
        mesh.close()


def run_plain(ctx, reqs):
    """Sequential FP64 answers; returns (rcs list, latency list in ms, wall s)."""
    rcs, lat = [], []
    ctx.sync()
    t0 = time.perf_counter()
    for r in reqs:
        t1 = time.perf_counter()
        rcs.append(solve_plain(ctx, r))
        lat.append((time.perf_counter() - t1) * 1e3)
    return rcs, lat, time.perf_counter() - t0


def run_served(srv, reqs, precision, rate=None, seed=SEED):
    """Submit `reqs` to `srv` (all at t=0 when rate is None, else Poisson
    arrivals at `rate` requests/s) and collect every result.  Returns
    (results in request order, server stats)."""
    srv.reset_stats()
    ids = [None] * len(reqs)
    arrivals = None
    if rate is not None:
        gaps = np.random.default_rng(seed + 1).exponential(1.0 / rate, len(reqs))
        arrivals = np.cumsum(gaps) - gaps[0]

    def feed():
        t0 = time.perf_counter()
        # all at t=0: the whole stream is queued before any worker takes a
        # request, so that the server orders it as one queue
        if arrivals is None:
            srv.hold(True)
        try:
            for i, r in enumerate(reqs):
                if arrivals is not None:
                    dt = arrivals[i] - (time.perf_counter() - t0)
                    if dt > 0:
                        time.sleep(dt)
                ids[i] = srv.submit(r["angles"], kind=r["kind"], subdiv=r["subdiv"], axes=r["axes"],
                                    formulation=r["formulation"], alpha=r["alpha"], freq=r["freq"],
                                    precision=precision)
        finally:
            if arrivals is None:
                srv.hold(False)

    th = threading.Thread(target=feed)
    th.start()
    got = {}
    while len(got) < len(reqs):
        res = srv.wait(-1)
        if res is None:
            if not th.is_alive() and len(got) + 0 >= sum(i is not None for i in ids):
                break
            time.sleep(0.001)
            continue
        got[res["id"]] = res
    th.join()
    srv.drain()
    by_id = {i: k for k, i in enumerate(ids)}
    out = [None] * len(reqs)
    for rid, res in got.items():
        out[by_id[rid]] = res
    return out, srv.stats()
