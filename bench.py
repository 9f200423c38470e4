"""Benchmark of the MoM hot path on one MI355X.

Workload (default): CFIE (alpha = 0.5) on a PEC sphere r = 1 m at 600 MHz,
icosphere subdivided 5 times (20480 triangles, 30720 RWG unknowns), 181
monostatic right-hand sides (theta 0..180 deg, 1 deg step, phi 0, pole 0).

One step = one complete solve on the FP32+refine path:
  complex-FP32 fill (3x3 Gauss, FP64 self terms) -> 181 RHS columns ->
  FP32 LU (MFMA trailing updates) -> FP64 iterative refinement (FP64 MFMA
  residual products) -> monostatic RCS of every column.
`value` = solves per second over the timed steps.  The plain FP64 path
(FP64 fill, FP64 LU, FP64 triangular solves) is timed once beside it, and
the largest RCS difference between the two paths is reported.

Roofline: the dominant kernel is the LU trailing update (MFMA GEMM); its
algorithmic flops per launch are 8*M*Nc*K (complex update of an M x Nc block
with rank K), timed with HIP events on the library's stream.
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# The solve server runs two streams per worker context; HIP maps streams onto
# GPU_MAX_HW_QUEUES hardware queues (4 by default, and environments may
# export 4 explicitly), and streams that share a queue execute in order.
# main() writes --hw-queues into the environment before HIP initialises
# (8 workers on 4 queues measured 103 solves/s against 171 on 32).

PEAK = {"f32_mfma": 157.3, "f64_mfma": 78.6, "bf16_mfma": 2500.0, "hbm": 8000.0}
# HBM bytes per launch of the split trailing update on the 30720-unknown
# workload (nb 512), from the FETCH_SIZE / WRITE_SIZE counter runs recorded in
# profiles/lu_gemm_split_cfie_sphere_30720.md (fetch scaled by 2.0 after a 1 GiB
# copy calibration), beside the algorithmic bytes of the same launches
TRAFFIC_SPLIT_30720 = {"measured": 4.681e9, "algorithmic": 2.644e9}

CONFIGS = {
    # name: (kind, subdiv, radius, freq, alpha, quad, nrhs-angles)
    "c2": ("CFIE", 5, 1.0, 6e8, 0.5, 3, np.arange(0.0, 181.0, 1.0)),
    "c1": ("EFIE", 3, 0.5, 3e8, 1.0, 7, np.array([0.0])),
    "s4": ("CFIE", 4, 1.0, 6e8, 0.5, 3, np.arange(0.0, 181.0, 1.0)),
    "c3": None,  # 256 uneven EFIE ellipsoid problems (emfa_amd.workload.batch_problems)
    "c5": None,  # MLFMA CFIE, icosphere s=6 r=1 m at 1.5 GHz, GMRES(30) to 1e-4
    "c4": None,  # mixed serving stream (emfa_amd.workload.mixed_stream)
    "fastwt": None,  # IP-SIE (FastWT) on a nonconformal cube, 1.5 GHz, BiCGSTAB to 1e-4
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--hw-queues", type=int, default=32, choices=range(1, 33), metavar="1..32",
                    help="GPU_MAX_HW_QUEUES for this run (streams of the server's worker contexts)")
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default 3; c4: requests, default 64)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c2", choices=sorted(CONFIGS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--precondition", action="store_true",
                    help="fastwt: sparse approximate inverse (preconditioning: true, row/col 0.9 / 0.87)")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64-path comparison")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-stream summary of the default line")
    ap.add_argument("--no-tune", action="store_true", help="skip the LU autotuner (built-in nb / panel choice)")
    ap.add_argument("--pivoted-fp32-only", action="store_true",
                    help="keep the FP32 factor on FP32-MFMA updates with partial pivoting (the tuner does not "
                         "consider the bf16-split update or panels without interchanges)")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="copy a buffer of this many MiB once (a plain copy kernel of known size, to scale "
                         "hardware traffic counters)")
    ap.add_argument("--mixed-requests", type=int, default=64)
    ap.add_argument("--workers", type=int, default=8)
    ap.add_argument("--budget-gb", type=float, default=96.0)
    ap.add_argument("--no-sweep", action="store_true", help="skip the Poisson rate sweep of the serving legs")
    ap.add_argument("--no-fp32-update", action="store_true",
                    help="skip the extra factorisation with the FP32-MFMA trailing update")
    ap.add_argument("--no-library", action="store_true",
                    help="skip the vendor-GEMM comparison on the trailing-update shapes")
    ap.add_argument("--library-gemm", type=int, nargs=2, metavar=("N", "NB"), default=None,
                    help="(child process) time the vendor GEMM on the trailing-update shapes of an N-unknown LU")
    ap.add_argument("--served-passes", type=int, default=5,
                    help="saturated server passes over the request stream; the median is reported")
    return ap.parse_args()


def monostatic_dirs(theta_deg):
    th = np.radians(theta_deg)
    return np.stack([np.sin(th), 0 * th, np.cos(th)], 1)


def host_cpu():
    """What the host offers: logical CPUs, the ones this process may use, model."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "usable": len(os.sched_getaffinity(0)), "model": model}


def write_nas(path, xyz, tri):
    """NASTRAN in the dialect the reference's Mesh::loadNasFile reads
    (EMComponent/Mesh.cpp:261-309): 9 header lines, GRID* pairs, CTRIA3."""
    with open(path, "w") as f:
        for i in range(9):
            f.write(f"$ header line {i}\n")
        for i, p in enumerate(xyz):
            f.write(f"GRID*   {i + 1:8d}  {p[0]: .9e}  {p[1]: .9e}\n")
            f.write(f"*       {i + 1:8d}  {p[2]: .9e}\n")
        for i, t in enumerate(tri):
            f.write(f"CTRIA3 {i + 1} 1 {t[0] + 1} {t[1] + 1} {t[2] + 1}\n")
        f.write("ENDDATA\n")


def cpu_baseline_reference(kind, radius, freq, alpha, n, nrhs, sample_subdiv=3, reps=3, lu_n=2048):
    """The reference's own code, single-threaded as it runs, on this host.

    Fill, excitation and far field: oracle/_ref/emfa_ref (the reference's
    EFIE/CFIE sources compiled unmodified, complex<float>) times its own
    fillZ, one fillV and getBiRCS on an icosphere of the same radius,
    frequency and formulation with `sample_subdiv` subdivisions.  Every RWG
    pair costs the fill the same whatever the mesh size, so fillZ scales by
    (N / n)^2, fillV and the far field by N / n per right-hand side.
    Solve: LAPACK cgesv (the routine behind the reference's arma::solve,
    EMSolver/MFIE.cpp:64; here scipy's OpenBLAS limited to one thread) on a
    lu_n x lu_n matrix with the workload's right-hand sides, scaled by N^3.
    Each part is the median of `reps` runs.  Returns None when the
    reference binary has not been built."""
    import subprocess
    import tempfile
    import scipy.linalg.lapack as lp
    from threadpoolctl import threadpool_limits
    from emfa_amd import hip
    exe = os.path.join(ROOT, "oracle", "_ref", "emfa_ref")
    if not os.path.exists(exe):
        return None
    xyz, tri = hip.icosphere(sample_subdiv, radius)
    runs = []
    with tempfile.TemporaryDirectory() as d:
        nas = os.path.join(d, "sample.nas")
        write_nas(nas, xyz, tri)
        for _ in range(reps):
            r = subprocess.run([exe, "time", kind.lower(), nas, repr(float(freq)), repr(float(alpha))], cwd=d,
                               stdout=subprocess.DEVNULL, stderr=subprocess.PIPE, text=True, check=True)
            ns, tz, tv, trcs = r.stderr.split()[-4:]
            runs.append((int(ns), float(tz), float(tv), float(trcs)))
    ns = runs[0][0]
    tz, tv, trcs = (float(np.median([x[i] for x in runs])) for i in (1, 2, 3))
    fill_s = tz * (n / ns) ** 2
    rhs_s = tv * (n / ns) * nrhs
    rcs_s = trcs * (n / ns) * nrhs
    lu_n = min(lu_n, n)
    rng = np.random.default_rng(0)
    A0 = (rng.standard_normal((lu_n, lu_n)) + 1j * rng.standard_normal((lu_n, lu_n))).astype(np.complex64)
    B0 = (rng.standard_normal((lu_n, nrhs)) + 1j * rng.standard_normal((lu_n, nrhs))).astype(np.complex64)
    tl = []
    with threadpool_limits(limits=1):
        for _ in range(reps):
            A, B = A0.copy(), B0.copy()
            t0 = time.perf_counter()
            lp.cgesv(A, B, overwrite_a=1, overwrite_b=1)
            tl.append(time.perf_counter() - t0)
    lu_s = float(np.median(tl)) * (n / lu_n) ** 3
    total = fill_s + rhs_s + lu_s + rcs_s
    return {
        "value": 1.0 / total,
        "unit": "solves/s",
        "cores": 1,
        "kind": "reference",
        "sample": (f"reference {kind} code compiled from its sources (oracle/_ref/emfa_ref, complex<float>, one "
                   f"thread): fillZ on icosphere s={sample_subdiv} ({ns} RWG) {tz:.2f}s -> {fill_s:.0f}s at "
                   f"{n} RWG by (N/n)^2; fillV {tv * 1e3:.2f} ms and getBiRCS {trcs * 1e3:.2f} ms per direction "
                   f"scaled by N/n x {nrhs}; LAPACK cgesv one thread {lu_n}x{lu_n} with {nrhs} RHS "
                   f"{np.median(tl):.2f}s -> {lu_s:.0f}s by N^3; medians of {reps} runs; est. full solve "
                   f"{total:.0f}s"),
        "phases_s": {"fill": fill_s, "rhs": rhs_s, "lu_solve": lu_s, "rcs": rcs_s},
        "host": host_cpu(),
    }


def mlfma_cpu_baseline(radius, freq, alpha, near_nnz, n, reps=3, sample_subdiv=3):
    """A lower bound on the reference's CPU time for one MLFMA solve: its
    near-field matrix alone.  MLFMA::calculateNearZ (EMSolver/MLFMA.cpp:
    1004-1131) evaluates the CFIE kernels of Miscellaneous.cpp for every near
    pair, the cost per pair that the reference's own CFIE::fillZ shows; that
    cost is timed here (oracle/_ref/emfa_ref, one thread, median of `reps`)
    on a sample icosphere and multiplied by the operator's near entries.
    Far-field setup, the translations and the iterations are left out, so the
    rate is an upper bound on the reference's solves per second.  The
    reference's MLFMA.cpp itself does not compile here (an Armadillo
    constructor at MLFMA.cpp:1033,1317)."""
    import subprocess
    import tempfile
    from emfa_amd import hip
    exe = os.path.join(ROOT, "oracle", "_ref", "emfa_ref")
    if not os.path.exists(exe):
        return None
    xyz, tri = hip.icosphere(sample_subdiv, radius)
    runs = []
    with tempfile.TemporaryDirectory() as d:
        nas = os.path.join(d, "sample.nas")
        write_nas(nas, xyz, tri)
        for _ in range(reps):
            r = subprocess.run([exe, "time", "cfie", nas, repr(float(freq)), repr(float(alpha))], cwd=d,
                               stdout=subprocess.DEVNULL, stderr=subprocess.PIPE, text=True, check=True)
            ns, tz = r.stderr.split()[-4:-2]
            runs.append(float(tz) / float(ns) ** 2)
    per_pair = float(np.median(runs))
    near_s = per_pair * near_nnz
    return {
        "value": 1.0 / near_s, "unit": "solves/s", "cores": 1, "kind": "reference",
        "sample": (f"reference CFIE kernels (oracle/_ref/emfa_ref, one thread): {per_pair * 1e6:.2f} us per RWG pair "
                   f"from CFIE::fillZ on icosphere s={sample_subdiv} at the same frequency (median of {reps}) x "
                   f"{near_nnz} near entries = {near_s:.0f}s for MLFMA::calculateNearZ alone; far-field setup, "
                   f"translations and iterations not included (an upper bound on the reference's rate)"),
        "host": host_cpu(),
    }


def cpu_baseline(xyz, tri, rwg, kind, k, alpha, quad, n_rhs, budget_s):
    """Parallel port: the oracle's FP64 fill (the reference's RWG-pair loop,
    OpenMP over test functions on every host core) on a row sample + LAPACK
    cgesv on all cores on a smaller matrix, scaled to one full solve."""
    from oracle import oracle as O
    import scipy.linalg.lapack as lp
    n = rwg.shape[0]
    threads = O.max_threads()
    # fill: grow the row sample until ~40% of the budget is spent
    rows = 8
    while True:
        t0 = time.perf_counter()
        if kind == "CFIE":
            O.fill_cfie(xyz, tri, rwg, k, alpha, quad, 0, rows=(0, rows))
        else:
            O.fill_efie(xyz, tri, rwg, k, quad, rows=(0, rows))
        tf = time.perf_counter() - t0
        if tf > 0.4 * budget_s or rows >= n:
            break
        rows = min(n, rows * 2)
    fill_est = tf * n / rows
    # LU + solve: cgesv on ns x ns with the same RHS count, scaled by N^3
    ns = 1024
    while True:
        rng = np.random.default_rng(0)
        A = (rng.standard_normal((ns, ns)) + 1j * rng.standard_normal((ns, ns))).astype(np.complex64)
        B = (rng.standard_normal((ns, n_rhs)) + 1j * rng.standard_normal((ns, n_rhs))).astype(np.complex64)
        t0 = time.perf_counter()
        lp.cgesv(A, B, overwrite_a=1, overwrite_b=1)
        tl = time.perf_counter() - t0
        if tl > 0.3 * budget_s or ns >= n:
            break
        ns = min(n, ns * 2)
    lu_est = tl * (n / ns) ** 3
    total = fill_est + lu_est
    return {
        "value": 1.0 / total,
        "unit": "solves/s",
        "cores": int(threads),
        "kind": "port",
        "host": host_cpu(),
        "sample": (f"oracle {kind} fill (FP64, RWG-pair loop as EFIE.cpp:603-696/CFIE.cpp:485-568) on rows 0..{rows} "
                   f"of {n} ({tf:.1f}s -> {fill_est:.0f}s for all rows) + LAPACK cgesv (arma::solve's routine) "
                   f"{ns}x{ns} with {n_rhs} RHS ({tl:.1f}s -> {lu_est:.0f}s scaled by N^3); est. full solve "
                   f"{total:.0f}s"),
    }


RATE_SWEEP = (0.25, 0.5, 1.0, 2.0, 4.0, 8.0)


def stream_leg(reqs, workers, budget, latency_rate=0.5, precision=None, passes=5, rates=RATE_SWEEP):
    """Answer `reqs` on the plain FP64 path (sequential, nothing cached) and
    through the solve server (FP32+refine, `workers` contexts, resident
    factors): `passes` times with every request queued at t=0 (capacity: the
    median pass; one pass varies by about +-3 %), then once on Poisson
    arrivals at `latency_rate` x that capacity (latency), each pass on a
    fresh server.  Every pass's RCS is checked.  Returns a summary dict."""
    from emfa_amd import hip, serve, workload
    precision = hip.F32 if precision is None else precision
    # Kernel code objects load on first use: both paths are warmed on the
    # sizes they will meet before anything is timed (the plain path on the
    # smallest and the largest request, the server on one untimed pass over
    # the stream with a server that is then discarded, so no factor stays).
    ctx = hip.Context(0)
    by_size = sorted(reqs, key=workload.unknowns)
    for r in (by_size[0], by_size[-1]):
        workload.solve_plain(ctx, r)
    plain, plain_lat, plain_wall = workload.run_plain(ctx, reqs)
    ctx.close()
    with serve.Server(workers=workers, hbm_budget=budget) as srv:
        workload.run_served(srv, reqs, precision)
    runs = []
    for _ in range(max(1, passes)):
        with serve.Server(workers=workers, hbm_budget=budget) as srv:
            runs.append(workload.run_served(srv, reqs, precision))
    per_pass = [st_["solves_per_s"] for _, st_ in runs]
    st = runs[sorted(range(len(runs)), key=lambda i: per_pass[i])[len(runs) // 2]][1]
    cap = st["solves_per_s"]
    with serve.Server(workers=workers, hbm_budget=budget) as srv:
        lat, st2 = workload.run_served(srv, reqs, precision, rate=latency_rate * cap)
    err = max(float(np.abs(p - r["rcs_db"]).max()) for out, _ in runs for p, r in zip(plain, out))
    err2 = max(float(np.abs(p - r["rcs_db"]).max()) for p, r in zip(plain, lat))
    # Poisson arrivals at multiples of the measured capacity, each on a fresh server
    sweep = []
    for f in rates:
        with serve.Server(workers=workers, hbm_budget=budget) as srv:
            out_r, st_r = workload.run_served(srv, reqs, precision, rate=f * cap)
        err2 = max(err2, max(float(np.abs(p - r["rcs_db"]).max()) for p, r in zip(plain, out_r)))
        sweep.append({"rate_x_capacity": f, "arrival_rate_per_s": f * cap, "p50_ms": st_r["p50_ms"],
                      "p99_ms": st_r["p99_ms"], "mean_ms": st_r["mean_ms"], "solves_per_s": st_r["solves_per_s"]})
    fb = max(int(sum(r["fp64_fallback"] for r in out)) for out, _ in runs)
    errors = sorted({r["error"] for out, _ in runs for r in out if r.get("status", 0) != 0} |
                    {r["error"] for r in lat if r.get("status", 0) != 0})
    ns = [workload.unknowns(r) for r in reqs]
    pl = np.sort(plain_lat)
    return {
        "requests": len(reqs), "unknowns_min": int(min(ns)), "unknowns_max": int(max(ns)),
        "rhs_total": int(sum(len(r["angles"]) for r in reqs)),
        "workers": workers, "hbm_budget_gb": budget / 2**30,
        "served_solves_per_s": cap, "served_passes_solves_per_s": per_pass,
        "served_rhs_per_s": st["rhs_per_s"], "served_wall_s": st["wall_s"],
        "served_cache_hits": st["cache_hits"], "served_failed": max(st_["failed"] for _, st_ in runs),
        "served_errors": errors[:5],
        "served_fp64_fallbacks": fb,
        "served_pivot_fallbacks": int(st.get("pivot_fallbacks", 0)),
        "served_batched_requests": int(st.get("batched", 0)),
        "served_factor_batches": int(st.get("factor_batches", 0)),
        "served_factors_in_batches": int(st.get("batch_factors", 0)),
        "served_peak_gb": st["peak_bytes"] / 2**30,
        "plain_fp64_solves_per_s": len(reqs) / plain_wall, "plain_fp64_wall_s": plain_wall,
        "plain_fp64_p50_ms": float(np.percentile(pl, 50)), "plain_fp64_p99_ms": float(np.percentile(pl, 99)),
        "speedup_vs_plain_fp64": cap / (len(reqs) / plain_wall),
        "latency_run": {"arrival_rate_per_s": latency_rate * cap, "p50_ms": st2["p50_ms"], "p99_ms": st2["p99_ms"],
                        "mean_ms": st2["mean_ms"], "max_ms": st2["max_ms"], "solves_per_s": st2["solves_per_s"]},
        "rate_sweep": sweep,
        # the median capacity pass, per request: unknowns, RHS, factor reused, batch size, service ms
        "served_requests": [[workload.unknowns(q), len(q["angles"]), int(bool(o["cache_hit"])), int(o["batch_size"]),
                             round(float(o["service_ms"]), 2)]
                            for q, o in zip(reqs, runs[sorted(range(len(runs)), key=lambda i: per_pass[i])[len(runs) // 2]][0])],
        "plain_requests_ms": [round(float(x), 2) for x in plain_lat],
        "saturated_p50_ms": st["p50_ms"], "saturated_p99_ms": st["p99_ms"],
        "rcs_max_abs_diff_db_vs_fp64": max(err, err2),
    }


def library_gemm_ceiling(n, nb):
    """Runs library_gemm_child in a child process (torch carries its own HIP
    runtime, which cannot initialise beside this process's) and returns its
    result."""
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--library-gemm", str(n), str(nb)],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=600)
    lines = [ln for ln in r.stdout.splitlines() if ln.startswith("{")]
    if r.returncode != 0 or not lines:
        return {"failed": r.stdout[-300:]}
    return json.loads(lines[-1])


def library_gemm_child(n, nb):
    """The outer trailing-update products of one n-unknown LU (M = n - (i+1) nb
    rows, 2 M real columns, K = 2 nb reals for i = 0, 1, ...) through the vendor
    GEMM (torch.mm, hipBLASLt on ROCm), timed on torch's stream: bf16 operands
    with the three split products stacked along K (one product with K = 6 nb,
    FP32 output, no C read), and the plain FP32 product.  Returns the total ms
    and the algorithmic TFLOP/s (8 M^2 nb per update, counted once) of each,
    for comparison with the build's kernels on the same shapes."""
    import torch
    Ms = [n - (i + 1) * nb for i in range(n // nb + 1) if n - (i + 1) * nb > 0]
    if not Ms:
        return None
    Kr, Mx = 2 * nb, Ms[0]
    flops = sum(8.0 * M * M * nb for M in Ms)
    dev = torch.device("cuda")
    out = {"updates": len(Ms), "nb": nb, "shape_first": [Mx, 2 * Mx, Kr]}
    for name, dt, kmul in (("bf16x3", torch.bfloat16, 3), ("fp32", torch.float32, 1)):
        A = torch.randn(Mx, kmul * Kr, device=dev).to(dt)
        B = torch.randn(2 * Mx, kmul * Kr, device=dev).to(dt)
        kw = {"out_dtype": torch.float32} if dt == torch.bfloat16 else {}

        def run():
            for M in Ms:
                torch.mm(A[:M], B[:2 * M].t(), **kw)
        run()  # warm (library heuristics, allocator)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        run()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1)
        tf = flops / (ms * 1e-3) / 1e12
        peak = PEAK["bf16_mfma"] if dt == torch.bfloat16 else PEAK["f32_mfma"]
        out[name] = {"ms": ms, "achieved": tf, "frac": tf / peak, "peak": peak}
        del A, B
        torch.cuda.empty_cache()
    return out


def main():
    a = parse()
    if a.library_gemm:
        print(json.dumps(library_gemm_child(*a.library_gemm)))
        return
    if a.gpus != 1:
        print(f"bench.py measures one GPU only: this build has no multi-GPU path (asked for --gpus {a.gpus}).",
              file=sys.stderr)
        sys.exit(2)
    # Nothing has touched HIP yet: the runtime reads this when it initialises.
    os.environ["GPU_MAX_HW_QUEUES"] = str(a.hw_queues)
    from emfa_amd import hip
    if a.config in ("c3", "c4"):
        if a.steps is None:
            a.steps = 64
        return main_stream(a)
    if a.config == "c5":
        if a.steps is None:
            a.steps = 2
        return main_mlfma(a)
    if a.config == "fastwt":
        if a.steps is None:
            a.steps = 2
        return main_fastwt(a)
    if a.steps is None:
        a.steps = 3

    kind_s, subdiv, radius, freq, alpha, quad, thetas = CONFIGS[a.config]
    kind = hip.CFIE if kind_s == "CFIE" else hip.EFIE
    xyz, tri = hip.icosphere(subdiv, radius)
    rwg, _ = hip.rwg_build(tri)
    n = rwg.shape[0]
    k = hip.wavenumber(freq)
    inc = hip.inc_table([(t, 0.0, 0.0) for t in thetas])
    nrhs = inc.shape[0]
    dirs = monostatic_dirs(thetas)
    cols = np.arange(nrhs, dtype=np.int32)

    ctx = hip.Context(0)
    if a.calib_copy_mb > 0:
        rows = a.calib_copy_mb * 2**20 // 16
        Ca, Cb = hip.Mat(ctx, rows, 1, hip.F64), hip.Mat(ctx, rows, 1, hip.F64)
        ctx.sync()
        Cb.copy_from(Ca)
        ctx.sync()
        Ca.close()
        Cb.close()
    # before the timed region; the FP32 path may pick the bf16-split update
    tuning = None if a.no_tune else ctx.autotune_lu(n, hip.F32, for_refinement=not a.pivoted_fp32_only)
    mesh = hip.Mesh(ctx, xyz, tri, rwg)
    Z = hip.Mat(ctx, n, n, hip.F32)
    B = hip.Mat(ctx, n, nrhs, hip.F64)
    X = hip.Mat(ctx, n, nrhs, hip.F64)
    reports = []

    def step():
        hip.fill(ctx, mesh, Z, kind, k, quad, alpha=alpha)
        hip.rhs_fill(ctx, mesh, B, kind, k, alpha, inc)
        lu = hip.LU(ctx, Z, hip.F32)
        rep, ok = hip.refine(ctx, Z, lu, B, X, tol=1e-10, maxit=8)
        reports.append((rep.iterations, rep.relres_first, rep.relres_final, ok))
        r = hip.rcs(ctx, mesh, k, X, dirs, cols)
        lu.close()
        return r

    for _ in range(a.warmup):
        step()
    ctx.sync()
    ctx.prof_reset()
    ctx.prof(True)
    # each step ends in a device-to-host copy of its RCS values, so the host
    # clock around one step is that step's latency
    step_s = []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        ts = time.perf_counter()
        r32 = step()
        step_s.append(time.perf_counter() - ts)
    ctx.sync()
    dt = (time.perf_counter() - t0) / a.steps
    fams = {}
    for f in ["fill_pairs", "fill_gather", "rhs", "lu_panel", "lu_laswp", "lu_trsm", "lu_gemm", "lu_split",
              "lu_gemm_split", "lu_dinv", "solve", "solve_trsm", "solve_apply", "solve_gemm", "refine_gemm", "rcs"]:
        ms, cnt, work = ctx.prof_get(f)
        fams[f] = {"ms_per_step": ms / a.steps, "launches_per_step": cnt / a.steps, "work_per_step": work / a.steps}
    ctx.prof(False)
    # the FP32 fill: k_fill_pairs (T^2 Q^2 Green's evaluations, writing the
    # per-triangle-pair scratch of 9 complex values per pair) and k_fill_gather
    # (per side of every RWG: two scratch reads and a read-modify-write of Z
    # per entry), beside the N^2 complex entries of Z itself
    T = int(tri.shape[0])
    pairs_ms, gather_ms = fams["fill_pairs"]["ms_per_step"], fams["fill_gather"]["ms_per_step"]
    scratch_b = 9.0 * T * T * 8
    gather_b = 2 * (2 * 8.0 + 2 * 8.0) * n * n
    z_b = 8.0 * n * n
    fill_roof = {
        "pairs": {"ms": pairs_ms, "green_evals": fams["fill_pairs"]["work_per_step"],
                  "evals_per_s": fams["fill_pairs"]["work_per_step"] / (pairs_ms * 1e-3) if pairs_ms > 0 else 0.0,
                  "scratch_write_bytes": scratch_b,
                  "write_gb_s": scratch_b / (pairs_ms * 1e-3) / 1e9 if pairs_ms > 0 else 0.0,
                  "bound": "valu (sqrt, reciprocal, sin/cos per evaluation)"},
        "gather": {"ms": gather_ms, "bytes": gather_b,
                   "gb_s": gather_b / (gather_ms * 1e-3) / 1e9 if gather_ms > 0 else 0.0,
                   "frac_hbm": gather_b / (gather_ms * 1e-3) / 1e9 / PEAK["hbm"] if gather_ms > 0 else 0.0,
                   "bound": "hbm"},
        "z_bytes": z_b,
        "traffic_over_z": (scratch_b + gather_b + z_b) / z_b,
        "note": "bytes as the kernels move them (scratch write, two scratch reads and a read-modify-write of Z "
                "per entry and side, the zeroing of Z), not counter readings",
    }
    # the dominant kernel: the trailing update of the outer panels -- bf16-split
    # planes (k_gemm_bf3r, bf16 MFMA) when the tuner chose them, else FP32 MFMA
    split = fams["lu_gemm_split"]["launches_per_step"] > 0
    g = fams["lu_gemm_split" if split else "lu_gemm"]
    gemm_tflops = g["work_per_step"] / (g["ms_per_step"] * 1e-3) / 1e12 if g["ms_per_step"] > 0 else 0.0
    avg_launch_ms = g["ms_per_step"] / max(g["launches_per_step"], 1)
    if split:
        roof = {"bound": "mfma", "kernel": "lu_gemm_split (k_gemm_bf3r, V_MFMA_F32_32X32X16_BF16, 3 products)",
                "achieved": gemm_tflops, "peak": PEAK["bf16_mfma"], "unit": "TFLOP/s",
                "frac": gemm_tflops / PEAK["bf16_mfma"],
                "traffic": TRAFFIC_SPLIT_30720["measured"] if n == 30720 else None,
                "traffic_algorithmic": TRAFFIC_SPLIT_30720["algorithmic"] if n == 30720 else None,
                "note": "achieved counts 8 M N K complex-FP32 flops once; the kernel issues three bf16 MFMA "
                        "products per flop, so MFMA-pipe utilisation is 3 x frac; traffic in bytes per launch "
                        "from the counter runs of the same command (nb 512)",
                "avg_launch_ms": avg_launch_ms, "launches_per_step": g["launches_per_step"]}
    else:
        roof = {"bound": "mfma", "kernel": "lu_gemm (k_gemm_f32, V_MFMA_F32_32X32X2_F32)",
                "achieved": gemm_tflops, "peak": PEAK["f32_mfma"], "unit": "TFLOP/s",
                "frac": gemm_tflops / PEAK["f32_mfma"], "traffic": None,
                "avg_launch_ms": avg_launch_ms, "launches_per_step": g["launches_per_step"]}

    # the same factorisation with the FP32-MFMA trailing update (k_gemm_f32),
    # and the vendor GEMM on the same update shapes
    nb_used = tuning["nb"] if tuning else int(os.environ.get("EMFA_LU_NB", "256") or 256)
    fp32_update = None
    if split and not a.no_fp32_update:
        ctx.lu_update_mode(0)
        hip.fill(ctx, mesh, Z, kind, k, quad, alpha=alpha)
        ctx.sync()
        ctx.prof_reset()
        ctx.prof(True)
        lu0 = hip.LU(ctx, Z, hip.F32)
        ctx.sync()
        ms0, cnt0, work0 = ctx.prof_get("lu_gemm")
        ctx.prof(False)
        lu0.close()
        ctx.lu_update_mode(1)
        tf0 = work0 / (ms0 * 1e-3) / 1e12 if ms0 > 0 else 0.0
        fp32_update = {"bound": "mfma", "kernel": "lu_gemm (k_gemm_f32, V_MFMA_F32_32X32X2_F32), same LU with the "
                                                  "FP32 trailing update",
                       "achieved": tf0, "peak": PEAK["f32_mfma"], "unit": "TFLOP/s", "frac": tf0 / PEAK["f32_mfma"],
                       "avg_launch_ms": ms0 / max(cnt0, 1), "launches": cnt0, "ms": ms0}
    library = None
    if not a.no_library:
        try:
            library = library_gemm_ceiling(n, nb_used)
        except Exception as e:  # a comparison, never required
            library = {"failed": str(e)[:200]}
    roof["library_same_shapes"] = library

    extra = {}
    if not a.no_fp64:
        Z.close()
        ctx.lu_update_mode(0)
        tuning64 = None if a.no_tune else ctx.autotune_lu(n, hip.F64)
        Z64 = hip.Mat(ctx, n, n, hip.F64)
        B64 = hip.Mat(ctx, n, nrhs, hip.F64)
        ctx.sync()
        t1 = time.perf_counter()
        hip.fill(ctx, mesh, Z64, kind, k, quad, alpha=alpha)
        hip.rhs_fill(ctx, mesh, B64, kind, k, alpha, inc)
        lu64 = hip.LU(ctx, Z64, hip.F64)
        lu64.solve(B64)
        r64 = hip.rcs(ctx, mesh, k, B64, dirs, cols)
        ctx.sync()
        t64 = time.perf_counter() - t1
        extra = {"fp64_path_solves_per_s": 1.0 / t64, "fp64_path_s": t64, "fp64_lu_tuning": tuning64,
                 "speedup_vs_fp64_path": t64 / dt,
                 "rcs_max_abs_diff_db_vs_fp64_path": float(np.abs(r32 - r64).max())}
        lu64.close()
        Z64.close()
        B64.close()

    # the serving leg runs before the CPU baseline: the baseline's BLAS and
    # OpenMP threads keep spinning for a while and would take the host cores
    # that the server's worker threads launch from
    for m_ in (B, X, mesh):
        m_.close()
    mixed = None
    if not a.no_mixed:
        from emfa_amd import workload
        mixed = stream_leg(workload.mixed_stream(a.mixed_requests), a.workers, int(a.budget_gb * 2**30),
                           passes=a.served_passes, rates=() if a.no_sweep else RATE_SWEEP)

    cpu = cpu_par = None
    if not a.no_cpu:
        try:
            cpu = cpu_baseline_reference(kind_s, radius, freq, alpha, n, nrhs)
        except Exception as e:  # the baseline is reported, never required
            cpu = {"value": None, "unit": "solves/s", "cores": 1, "kind": "reference", "sample": f"failed: {e}"}
        try:
            cpu_par = cpu_baseline(xyz, tri, rwg, kind_s, k, alpha, quad, nrhs, a.cpu_seconds)
        except Exception as e:
            cpu_par = {"value": None, "unit": "solves/s", "cores": None, "kind": "port", "sample": f"failed: {e}"}
        if cpu is None:
            cpu, cpu_par = cpu_par, None

    out = {
        "metric": "solves/sec (FP32+refine path) + p50/p99 latency",
        "value": 1.0 / dt,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("c64 fill/LU (trailing update: 3 bf16 MFMA products, f32 accumulate) + f64 refinement"
                  if split else "c64 fill/LU + f64 refinement"),
        "data": "synthetic icosphere mesh",
        "config": {"workload": f"{kind_s} PEC sphere r={radius} m f={freq / 1e6:.0f} MHz icosphere s={subdiv}",
                   "unknowns": n, "triangles": int(tri.shape[0]), "rhs": nrhs, "quadrature": f"{quad}x{quad}",
                   "parallelism": "single GPU"},
        "latency_ms": {"p50": float(np.percentile(step_s, 50)) * 1e3, "p99": float(np.percentile(step_s, 99)) * 1e3,
                       "min": min(step_s) * 1e3, "max": max(step_s) * 1e3,
                       "note": f"percentiles over the {a.steps} timed steps (one solve each, back to back)"},
        "roofline": roof,
        "roofline_fp32_update": fp32_update,
        "phases_ms_per_step": {f: round(v["ms_per_step"], 3) for f, v in fams.items()},
        "fill": fill_roof,
        "lu_tuning": tuning,
        "refine": {"iterations": reports[-1][0], "relres_first": reports[-1][1], "relres_final": reports[-1][2],
                   "converged": reports[-1][3]},
        "cpu_baseline": cpu,
        "cpu_baseline_parallel": cpu_par,
    }
    out.update(extra)
    if mixed is not None:
        out["mixed_stream"] = mixed
    print(json.dumps(out))


def main_mlfma(a, subdiv=6, radius=1.0, freq=1.5e9, alpha=0.5, threshold=0.25, tol=1e-4, restart=30):
    """c5: one step = MLFMA setup (tree, tables, near matrix, patterns) +
    CFIE excitation + GMRES(30) to 1e-4 + bistatic RCS at 361 angles."""
    from emfa_amd import hip
    xyz, tri = hip.icosphere(subdiv, radius)
    rwg, _ = hip.rwg_build(tri)
    n = rwg.shape[0]
    k = hip.wavenumber(freq)
    inc = hip.inc_table([(0.0, 0.0, 0.0)])
    th = np.radians(np.arange(0.0, 361.0, 1.0))
    dirs = np.stack([np.sin(th), 0 * th, np.cos(th)], 1)
    ctx = hip.Context(0)
    mesh = hip.Mesh(ctx, xyz, tri, rwg)
    B = hip.Mat(ctx, n, 1, hip.F64)
    X = hip.Mat(ctx, n, 1, hip.F64)
    Z0 = hip.Mat(ctx, n, 1, hip.F64)
    out = {}

    def step():
        t0 = time.perf_counter()
        op = hip.MLFMA(ctx, mesh, k, alpha, threshold)
        ctx.sync()
        t1 = time.perf_counter()
        hip.rhs_fill(ctx, mesh, B, hip.CFIE, k, alpha, inc)
        X.copy_from(Z0)
        rep, ok = op.gmres(B, X, m=restart, tol=tol, max_restart=40)
        r = hip.rcs(ctx, mesh, k, X, dirs, np.zeros(len(th), dtype=np.int32))
        ctx.sync()
        t2 = time.perf_counter()
        out.update(setup_s=t1 - t0, solve_s=t2 - t1, iterations=rep.iterations, restarts=rep.restarts,
                   relres=rep.relres, converged=ok,
                   info=op.info(), rcs=r)
        op.close()

    for _ in range(a.warmup):
        step()
    ctx.prof_reset()
    ctx.prof(True)
    t0 = time.perf_counter()
    for _ in range(a.steps):
        step()
    ctx.sync()
    dt = (time.perf_counter() - t0) / a.steps
    fams = {}
    for f in ["mlfma_setup", "mlfma_near", "mlfma_agg", "mlfma_up", "mlfma_down", "mlfma_out"]:
        ms, cnt, work = ctx.prof_get(f)
        fams[f] = {"ms_per_step": ms / a.steps, "launches_per_step": cnt / a.steps, "work_per_step": work / a.steps}
    ctx.prof(False)
    info = out["info"]
    nf = fams["mlfma_near"]
    near_gbs = nf["work_per_step"] / (nf["ms_per_step"] * 1e-3) / 1e9 if nf["ms_per_step"] > 0 else 0.0
    # dominant kernel: the downward pass (translation + anterpolation, k_ml_down)
    dn = fams["mlfma_down"]
    down_gbs = dn["work_per_step"] / (dn["ms_per_step"] * 1e-3) / 1e9 if dn["ms_per_step"] > 0 else 0.0
    # what its gathers move through the caches: each far pair reads one
    # translation entry (8 B) and one radiation pattern (24 B) per sample
    K = [m * 2 * m for m in info["multipoles"]]
    gather_b = sum(info["far_pairs"][l] * K[l] * 32.0 for l in range(2, info["levels"] + 1))
    down_launch_ms = dn["ms_per_step"] / max(dn["launches_per_step"], 1)
    matvecs = out["iterations"] + out["restarts"] + 1  # Arnoldi steps + restart residuals + final residual
    mv_ms = sum(fams[f]["ms_per_step"] for f in ["mlfma_near", "mlfma_agg", "mlfma_up", "mlfma_down", "mlfma_out"])
    res = {
        "metric": "solves/sec (MLFMA CFIE, GMRES to 1e-4)",
        "value": 1.0 / dt,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 near/far, f64 self terms and GMRES",
        "data": "synthetic icosphere mesh",
        "config": {"workload": f"MLFMA CFIE alpha={alpha} PEC sphere r={radius} m f={freq / 1e9:.1f} GHz "
                               f"icosphere s={subdiv}, box threshold {threshold} lambda, GMRES({restart}) to {tol}",
                   "unknowns": n, "triangles": int(tri.shape[0]), "parallelism": "single GPU"},
        "mlfma": {"levels": info["levels"], "boxes": info["boxes"], "multipoles": info["multipoles"],
                  "near_nnz": info["near_nnz"], "device_gb": info["bytes"] / 2**30,
                  "setup_s": out["setup_s"], "solve_s": out["solve_s"], "gmres_iterations": out["iterations"],
                  "relres": out["relres"], "converged": out["converged"],
                  "matvec_ms": mv_ms / max(matvecs, 1)},
        "roofline": {"bound": "hbm", "kernel": "mlfma_down (k_ml_down: far-box translation + anterpolation)",
                     "achieved": down_gbs, "peak": PEAK["hbm"], "unit": "GB/s", "frac": down_gbs / PEAK["hbm"],
                     "traffic": None,
                     "avg_launch_ms": down_launch_ms, "launches_per_step": dn["launches_per_step"],
                     "gather_bytes_per_matvec": gather_b,
                     "gather_gb_per_s": gather_b * dn["launches_per_step"] / max(matvecs, 1)
                     / max(info["levels"] - 1, 1) / (dn["ms_per_step"] * 1e-3) / 1e9 if dn["ms_per_step"] > 0 else 0,
                     "note": "achieved counts each level's patterns, translation rows in use and parent patterns "
                             "once (algorithmic bytes); the kernel reads far-pair patterns and translation entries "
                             "through L2 (gather_gb_per_s), which bounds it"},
        "near_spmv": {"achieved_gb_per_s": near_gbs, "frac": near_gbs / PEAK["hbm"]},
        "phases_ms_per_step": {f: round(v["ms_per_step"], 3) for f, v in fams.items()},
        "cpu_baseline": None if a.no_cpu else mlfma_cpu_baseline(radius, freq, alpha, info["near_nnz"], n),
    }
    print(json.dumps(res))


def fastwt_cpu_baseline(near_nnz, freq, alpha, beta, reps=3, sample_res=(4, 5, 4, 5, 4, 5)):
    """A lower bound on the CPU time of one IP-SIE solve: its near matrix
    alone.  FastWT::calculateNearZ (EMSolver/FastWT.cpp:689-784) evaluates
    for every near pair the CFIE kernel of Miscellaneous.cpp, two line
    integrals and the penalty; the restatement (oracle/fastwt.py, the C
    oracle's kernels and numpy, one thread) is timed per pair on every pair
    of a small nonconformal cube at the same frequency (median of `reps`)
    and multiplied by the operator's near entries.  The reference's
    FastWT.cpp does not compile here."""
    from emfa_amd import hip, workload
    from oracle import fastwt as W
    from oracle import oracle as O
    xyz, tri = workload.nonconformal_cube(sample_res, 0.6)
    c = hip.contour_build(xyz, tri)
    C = W.Contour(c.hrwg, c.length, c.line_ptr, c.line_tri, c.line_seg)
    k = hip.wavenumber(freq)
    threads = O.max_threads()
    O.set_threads(1)
    try:
        runs = []
        for _ in range(reps):
            t0 = time.perf_counter()
            W.dense_parts(xyz, tri, C, k, beta, alpha)
            runs.append((time.perf_counter() - t0) / C.n ** 2)
    finally:
        O.set_threads(threads)
    per_pair = float(np.median(runs))
    near_s = per_pair * near_nnz
    return {
        "value": 1.0 / near_s, "unit": "solves/s", "cores": 1, "kind": "port",
        "sample": (f"restated IP-SIE near entries (oracle/fastwt.py, one thread): {per_pair * 1e6:.2f} us per "
                   f"half-RWG pair over all {C.n}^2 pairs of a nonconformal cube (median of {reps}) x {near_nnz} "
                   f"near entries = {near_s:.0f}s for FastWT::calculateNearZ alone; far-field set-up, translations "
                   f"and iterations not included (an upper bound on a CPU rate)"),
        "host": host_cpu(),
    }


def main_fastwt(a, res=(40, 48, 40, 48, 40, 48), size=1.0, freq=1.5e9, alpha=0.5, threshold=0.25, tol=1e-4,
                stab_factor=0.1, avg_size=0.07503, row_threshold=0.9, col_threshold=0.87):
    """fastwt: one step = contour edges + IP-SIE operator set-up (tree,
    tables, near matrix, patterns) + excitation + BiCGSTAB to 1e-4 + bistatic
    RCS at 361 angles, on a cube whose faces are meshed separately with 40
    and 48 cells per edge (nonconformal seams)."""
    from emfa_amd import hip, workload
    xyz, tri = workload.nonconformal_cube(res, size)
    rwg, _ = hip.rwg_build(tri)
    k = hip.wavenumber(freq)
    beta = stab_factor * abs(np.log10(avg_size))
    inc = hip.inc_table([(0.0, 0.0, 0.0)])
    th = np.radians(np.arange(0.0, 361.0, 1.0))
    dirs = np.stack([np.sin(th), 0 * th, np.cos(th)], 1)
    ctx = hip.Context(0)
    mesh = hip.Mesh(ctx, xyz, tri, rwg)
    out = {}

    def step():
        t0 = time.perf_counter()
        c = hip.contour_build(xyz, tri)
        op = hip.FastWT(ctx, mesh, c, k, alpha, threshold, beta)
        if a.precondition:
            out["precond_nnz"] = op.precondition(mesh, row_threshold, col_threshold)
        ctx.sync()
        t1 = time.perf_counter()
        n = op.n
        B = hip.Mat(ctx, n, 1, hip.F64)
        X = hip.Mat(ctx, n, 1, hip.F64)
        op.rhs(B, inc)
        rep, ok = op.bicgstab(B, X, tol=tol, max_iter=2000, variant=0)
        r = op.rcs(X, dirs)
        ctx.sync()
        t2 = time.perf_counter()
        out.update(setup_s=t1 - t0, solve_s=t2 - t1, iterations=rep.iterations, relres=rep.relres, converged=ok,
                   info=op.info(), rcs=r, n=n, nonconformal=int((c.hrwg[:, 5] == 1).sum()))
        B.close()
        X.close()
        op.close()

    for _ in range(a.warmup):
        step()
    ctx.prof_reset()
    ctx.prof(True)
    t0 = time.perf_counter()
    for _ in range(a.steps):
        step()
    ctx.sync()
    dt = (time.perf_counter() - t0) / a.steps
    fams = {}
    for f in ["mlfma_setup", "mlfma_near", "mlfma_agg", "mlfma_up", "mlfma_down", "mlfma_out"]:
        ms, cnt, work = ctx.prof_get(f)
        fams[f] = {"ms_per_step": ms / a.steps, "launches_per_step": cnt / a.steps, "work_per_step": work / a.steps}
    ctx.prof(False)
    info, n = out["info"], out["n"]
    dn = fams["mlfma_down"]
    down_gbs = dn["work_per_step"] / (dn["ms_per_step"] * 1e-3) / 1e9 if dn["ms_per_step"] > 0 else 0.0
    nf = fams["mlfma_near"]
    near_gbs = nf["work_per_step"] / (nf["ms_per_step"] * 1e-3) / 1e9 if nf["ms_per_step"] > 0 else 0.0
    matvecs = 2 * out["iterations"] + 1
    mv_ms = sum(fams[f]["ms_per_step"] for f in ["mlfma_near", "mlfma_agg", "mlfma_up", "mlfma_down", "mlfma_out"])
    res = {
        "metric": "solves/sec (FastWT IP-SIE, BiCGSTAB to 1e-4)",
        "value": 1.0 / dt,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 near/far, f64 self terms and BiCGSTAB",
        "data": "synthetic nonconformal cube mesh",
        "config": {"workload": f"FastWT IP-SIE alpha={alpha} PEC cube {size} m, faces of {res[0]} / {res[1]} cells "
                               f"per edge meshed separately, f={freq / 1e9:.1f} GHz, box threshold {threshold} "
                               f"lambda, BiCGSTAB to {tol}" + (f", SAI preconditioner (row / col thresholds "
                                                              f"{row_threshold} / {col_threshold} leaf edges)"
                                                              if a.precondition else ""),
                   "unknowns": n, "nonconformal": out["nonconformal"], "triangles": int(tri.shape[0]),
                   "parallelism": "single GPU"},
        "fastwt": {"levels": info["levels"], "boxes": info["boxes"], "multipoles": info["multipoles"],
                   "near_nnz": info["near_nnz"], "device_gb": info["bytes"] / 2**30,
                   "setup_s": out["setup_s"], "solve_s": out["solve_s"], "bicgstab_iterations": out["iterations"],
                   "relres": out["relres"], "converged": out["converged"], "matvec_ms": mv_ms / max(matvecs, 1),
                   "precond_nnz": out.get("precond_nnz")},
        "roofline": {"bound": "hbm", "kernel": "mlfma_down (k_ml_down<4>: far-box translation + anterpolation)",
                     "achieved": down_gbs, "peak": PEAK["hbm"], "unit": "GB/s", "frac": down_gbs / PEAK["hbm"],
                     "traffic": None, "avg_launch_ms": dn["ms_per_step"] / max(dn["launches_per_step"], 1),
                     "launches_per_step": dn["launches_per_step"],
                     "note": "achieved counts each level's four-channel patterns, translation rows in use and "
                             "parent patterns once (algorithmic bytes)"},
        "near_spmv": {"achieved_gb_per_s": near_gbs, "frac": near_gbs / PEAK["hbm"]},
        "phases_ms_per_step": {f: round(v["ms_per_step"], 3) for f, v in fams.items()},
        "cpu_baseline": None if a.no_cpu else fastwt_cpu_baseline(info["near_nnz"], freq, alpha, beta),
    }
    print(json.dumps(res))


def main_stream(a):
    """c3: the 256 uneven problems; c4: the mixed serving stream with
    `--steps` requests.  A step is one request; value = served solves/s."""
    from emfa_amd import workload
    if a.config == "c3":
        reqs = workload.batch_problems()
        name = "256 PEC ellipsoids, semi-axes U[0.3,1.2] m, icosphere s=2..4, 100-600 MHz, EFIE, 1-64 RHS"
    else:
        reqs = workload.mixed_stream(a.steps)
        name = ("mixed stream: spheres/ellipsoids (s=3,4) and 1x1 m plates (sizes stratified over 500-15000 RWG), "
                "EFIE/CFIE, 1-64 RHS, 30% repeat meshes")
    leg = stream_leg(reqs, a.workers, int(a.budget_gb * 2**30), passes=a.served_passes,
                     rates=() if a.no_sweep else RATE_SWEEP)
    lr = leg["latency_run"]
    out = {
        "metric": "solves/sec (served FP32+refine) + p50/p99 latency",
        "value": leg["served_solves_per_s"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": len(reqs),
        "warmup": 1,
        "ms_per_step": 1e3 / leg["served_solves_per_s"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 fill/LU + f64 refinement",
        "data": f"synthetic meshes, seed {workload.SEED}",
        "config": {"workload": name, "parallelism": "single GPU", "workers": a.workers},
        "latency_ms": {"p50": lr["p50_ms"], "p99": lr["p99_ms"],
                       "note": f"Poisson arrivals at {lr['arrival_rate_per_s']:.2f}/s (0.5x capacity)"},
        "stream": leg,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
