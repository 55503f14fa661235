"""Benchmark: k-batched Kadanoff-Baym time stepping on one MI355X.

Workload (the reference's k-batched lattice): 1D tight-binding chain eps_k = -2 cos(k),
Nk = 128, norb = 1, beta = 10, ntau = 400, nt = 400, h = 0.025, SolveOrder 5, local
second-order self-energy Sigma = U^2 G_loc(t,t')^2 G_loc(t',t) from Bubble1/Bubble2 with an
interaction quench U: 0 -> 2 at t = 0, self-consistent to distance_norm2 < 1e-12 at every
timestep.

A "step" is one full propagation sweep: timesteps SolveOrder+1 .. nt (6..400) of all Nk
k-points, each timestep with its predictor and every corrector iteration it needs, from
the converged start (timesteps 0..5) that equilibrium() + bootstrap() leave behind.  A sweep
reads only timesteps it has already recomputed, so every sweep repeats the first one bit
for bit (the bench checks that).  `value` = Nk * 395 * steps / wall time of the timed
sweeps (all data resident in HBM).

Beside the headline, in the same invocation:
  * the plain per-k loop (cntrx_dyson_timestep once per k, per iteration) over the whole
    sweep, with its G_loc compared bit for bit with the batched one;
  * the host oracle (the restated reference algorithm, OpenMP over k) on a bounded prefix
    of the same workload, which is also the parity check of the GPU sweep;
  * the uneven batch (64 independent bath solves, size1 in {1,2,4,8}, nt in [200,1500],
    ntau 300) through the batch server, against the per-solve plain loop and the host;
  * the multi-orbital k-batch (norb 8) at its per-k shape on a k subset;
  * the serving sweep: Poisson arrivals of requests from the uneven distribution with nt up
    to 2000 into the batch server, admitted by HBM footprint against a 200 GB budget.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

NK, NORB, BETA, NTAU, NT, H, K, U = 128, 1, 10.0, 400, 400, 0.025, 5, 2.0
TOL, MAXITER = 1e-12, 50
PEAK_HBM_GBS = 8000.0
PEAK_FP64_TFLOPS = 78.6
FIRST = K + 1  # first timestep of a sweep


def kernel_bytes(t, ntau, nk):
    """algorithmic HBM bytes of one batched launch of each kernel at timestep t
    (complex128 = 16 B; shared Sigma counted once per launch)."""
    row = 16.0 * (ntau + 1)
    tri = 16.0 * (t + 1) * (t + 2) / 2
    col = 16.0 * (t + 1)
    tv = nk * (t + 2) * row + row + col                 # G^tv history + G^M + out; Sigma rows
    ret = tri + nk * 2 * col                            # Sigma^R triangle; H, out row
    return {
        "k_tv": tv,
        "k_retrec": ret,
        "k_tvret1": tv + ret,                               # the two in one launch
        "k_lesconv": (t + 1) * row + tri + nk * (row + 2 * col),  # Sigma^tv hist, Sigma^< tri
        "k_lesrec": tri + nk * 2 * col,                    # Sigma^R triangle; gles, out col
    }


PMC_SUMMARY = os.path.join(ROOT, "profiles", "lattice_nk128_nt400_pmc.md")


def pmc_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc summary of this
    workload (FETCH_SIZE + WRITE_SIZE, each scaled by a 256 MiB copy), and the window that
    summary covers; counters cannot be read inside the timed run itself"""
    val, window = None, None
    try:
        with open(PMC_SUMMARY) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    vals = dict(kv.split("=") for kv in line.split(":", 1)[1].split())
                    if kernel in vals:
                        val = float(vals[kernel])
                elif line.startswith("traffic_window:"):
                    window = line.split(":", 1)[1].strip()
    except OSError:
        pass
    return val, window


def log(msg):
    """progress line on stderr (the JSON result is the only line on stdout)"""
    print("[bench %7.1fs] %s" % (time.perf_counter() - T_START, msg), file=sys.stderr, flush=True)


T_START = time.perf_counter()


def host_info():
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count()
    return {"nproc": os.cpu_count(), "affinity_cpus": aff, "cpu_model": model}


def solve_bytes(size1, ntau, n_from, n_to):
    """algorithmic HBM bytes of dyson_timestep n = n_from..n_to on one solve with its own
    Sigma: Sigma^tv and G^tv histories, Sigma^R and Sigma^< triangles, complex128"""
    n = np.arange(n_from, n_to + 1, dtype=np.float64)
    return float(16.0 * size1 ** 2 * np.sum(2 * (n + 1) * (ntau + 1) + (n + 1) * (n + 2)))


def truncate(r, nt):
    """the same request stopped at timestep nt"""
    from nessi_amd.workloads import make_request
    return make_request(r["size1"], nt, r["ntau"], r["beta"], r["h"], r["H"][:nt + 2], r["Hbath"], r["V"],
                        mu=r["mu"], sig=r["sig"], SolveOrder=r["SolveOrder"], mat_method=r["mat_method"])


def plain_solve(r):
    """one request through the per-solve entry points (dyson_mat, dyson_start, then
    dyson_timestep once per step), the reference's plain caller loop"""
    import torch
    from nessi_amd import cntr
    s, nt, ntau = r["size1"], r["nt"], r["ntau"]
    G = cntr.herm_matrix(nt, ntau, s, r["sig"])
    Sg = cntr.herm_matrix(nt, ntau, s, r["sig"])
    cntr.green_from_H(Sg, r["mu"], r["Hbath"], r["beta"], r["h"])
    Sg._storage.mul_(r["V"] ** 2)
    H = cntr.function(nt, s)
    H.data.copy_(torch.from_numpy(r["H"]))
    cntr.dyson(G, r["mu"], H, Sg, r["beta"], r["h"], r["SolveOrder"], r["mat_method"])
    return G


def herm_rel_err(g, X, nstop, ntau, comps=("ret", "les", "tv", "mat")):
    """max-abs error / max-abs value per component over timesteps 0..nstop (and the
    Matsubara component)"""
    o = (nstop + 1) * (nstop + 2) // 2
    ntv = (nstop + 1) * (ntau + 1)
    errs = {}
    for c, n in (("ret", o), ("les", o), ("tv", ntv), ("mat", ntau + 1)):
        if c in comps:
            ref = getattr(X, c)[:n]
            errs[c] = float(np.abs(g[c][:n] - ref).max() / np.abs(ref).max())
    return errs


def run_uneven(args):
    """Uneven batch: 64 seeded independent solves submitted at once to the batch server
    (every live solve advances one timestep per tick through the batched entry point),
    timed from the first submission until the last solve is done; the plain per-solve loop
    on the same requests cut to nt = 200 (both legs timed end to end); parity of one
    request per size against the oracle; the oracle on the host cores."""
    import torch
    from nessi_amd import batch_server, workloads
    rs = workloads.uneven_batch(64, seed=2024)
    batch_server.autotune(not args.no_autotune)
    S = batch_server.BatchServer(200 << 30)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ids = [S.submit(r) for r in rs]
    S.drain()
    torch.cuda.synchronize()
    secs = time.perf_counter() - t0
    log("uneven batch: 64 solves in %.1f s (%d ticks)" % (secs, S.ticks))
    k = rs[0]["SolveOrder"]
    nbytes = sum(solve_bytes(r["size1"], r["ntau"], k + 1, r["nt"]) for r in rs)
    from oracle import oracle
    NPAR = args.uneven_parity_steps
    par = {}
    for sz in (1, 2, 4, 8):
        i = next(i for i, r in enumerate(rs) if r["size1"] == sz)
        g = S.result(ids[i]).to_numpy()
        X, _ = oracle.dyson_solve(rs[i], nt_stop=NPAR)
        par["size1=%d" % sz] = herm_rel_err(g, X, NPAR, rs[i]["ntau"])
    log("uneven batch: parity vs oracle done")
    timesteps = S.timesteps
    for rid in ids:
        S.release(rid)
    S.close()
    # same requests cut to nt = 200: batch server vs plain loop, end to end
    cut = [truncate(r, 200) for r in rs[:args.plain_solves]]
    steps_cut = sum(r["nt"] - k for r in cut)
    S = batch_server.BatchServer(64 << 30)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for r in cut:
        S.submit(r)
    S.drain()
    torch.cuda.synchronize()
    batched_cut = time.perf_counter() - t0
    S.close()
    log("uneven batch, nt 200: batched %.2f s" % batched_cut)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for r in cut:
        plain_solve(r)
    torch.cuda.synchronize()
    plain_cut = time.perf_counter() - t0
    log("uneven batch, nt 200: plain loop %.2f s" % plain_cut)
    out = {
        "workload": "64 independent bath solves, size1 uniform {1,2,4,8}, nt uniform [200,1500], "
                    "ntau 300, h 0.02, beta uniform [2,20], H0 + 0.2 sin(Omega t) H1, V 0.4, seed 2024",
        "path": "C++ batch server (cntrx_server_*): admission by footprint against a device pool, "
                "then one batched dyson_timestep call over every live solve per tick",
        "value": timesteps / secs, "unit": "solve-timesteps/s", "seconds": secs,
        "timesteps": timesteps, "algorithmic_bytes": nbytes, "effective_GBs": nbytes / secs / 1e9,
        "hbm_roofline_frac": nbytes / secs / 1e9 / PEAK_HBM_GBS,
        "autotuner": batch_server.autotune_decisions(),
        "batched_vs_plain_nt200": {"solves": len(cut), "timesteps": steps_cut,
                                   "batched_value": steps_cut / batched_cut,
                                   "plain_value": steps_cut / plain_cut,
                                   "speedup": plain_cut / batched_cut},
        "parity_vs_oracle": {"timesteps_checked": "0..%d and Matsubara" % NPAR, "max_rel_err": par,
                             "tolerance": 1e-12,
                             "pass": all(e <= 1e-12 for d in par.values() for e in d.values())},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_uneven(rs, args)
    return out


def cpu_uneven(rs, args):
    """the oracle on the host cores: one thread per request (ctypes releases the GIL),
    each request's full setup plus its first CPU_STEPS timesteps"""
    from concurrent.futures import ThreadPoolExecutor
    from oracle import oracle
    oracle.lib()
    thr = max(1, min(args.cpu_threads, host_info()["affinity_cpus"] or 1))
    sample = [r for r in rs if r["size1"] <= 4][:thr]
    nstop = rs[0]["SolveOrder"] + args.cpu_uneven_steps
    t0 = time.perf_counter()
    with ThreadPoolExecutor(thr) as ex:
        list(ex.map(lambda r: oracle.dyson_solve(r, nt_stop=nstop), sample))
    secs = time.perf_counter() - t0
    log("uneven batch: host oracle sample %.1f s" % secs)
    steps = len(sample) * args.cpu_uneven_steps
    return {"value": steps / secs, "unit": "solve-timesteps/s", "cores": thr, "kind": "port",
            "sample": "%d requests of the batch with size1 <= 4, setup (green_from_H, dyson_mat, "
                      "dyson_start) plus timesteps %d..%d each, one thread per request"
                      % (len(sample), rs[0]["SolveOrder"] + 1, nstop)}


def run_multiorbital(args):
    """Multi-orbital k-batch at the per-k shape of the 16x16-grid workload (norb 8, beta 20,
    ntau 600, h 0.01, Sigma = 0.09 G_loc, Peierls-shifted H(k,t)) on an nkx x nkx subset of
    the k-points, propagated to timestep mo_steps; the last 10 timesteps are timed."""
    import torch
    from nessi_amd import multiorbital as mo
    nkx, nt = args.mo_nkx, args.mo_steps
    log("multiorbital: %d k-points, norb 8, ntau 600, to timestep %d" % (nkx * nkx, nt))
    L = mo.MultiOrbitalLattice(nkx, 8, 20.0, 600, nt, 0.01, g2=0.09)
    t0 = time.perf_counter()
    mat_it = L.matsubara(1e-12, 50)
    torch.cuda.synchronize()
    t_mat = time.perf_counter() - t0
    start_it = L.start(1e-12, 50)
    first = max(L.k + 1, nt - 9)
    for n in range(L.k + 1, first):
        L.step(n, 1e-12, 50)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    iters, flops = 0, 0.0
    for n in range(first, nt + 1):
        it = L.step(n, 1e-12, 50)
        iters += it
        # algorithmic FP64 work of one dyson_timestep: E(n) block products of 8 S^3 flops,
        # E(n) = (ntau+1)^2 + 2(n+1)(ntau+1) + 2(n+1)^2, per k-point and corrector iteration
        e_n = (L.ntau + 1.0) ** 2 + 2.0 * (n + 1) * (L.ntau + 1) + 2.0 * (n + 1) ** 2
        flops += it * L.nk * e_n * 8.0 * 8 ** 3
    torch.cuda.synchronize()
    secs = time.perf_counter() - t0
    steps = nt + 1 - first
    log("multiorbital: %d timed steps in %.3f s" % (steps, secs))
    return {"workload": "multi-orbital k-batch, norb 8, beta 20, ntau 600, h 0.01, Sigma = 0.09 G_loc, "
                        "H(k,t) = H_rand + eps(k - A(t)), %dx%d of the 16x16 k-grid" % (nkx, nkx),
            "value": L.nk * steps / secs, "unit": "k-timesteps/s", "timed_timesteps": [first, nt],
            "corrector_iterations_timed": iters, "ms_per_step": 1e3 * secs / steps,
            "matsubara_seconds": t_mat, "matsubara_iterations": mat_it, "start_iterations": start_it,
            "fp64": {"algorithmic_flops": flops, "achieved_TFLOPs": flops / secs / 1e12,
                     "peak_TFLOPs": PEAK_FP64_TFLOPS, "frac": flops / secs / 1e12 / PEAK_FP64_TFLOPS,
                     "count": "8 size1^3 E(n) per k-point and corrector iteration, "
                              "E(n) = (ntau+1)^2 + 2(n+1)(ntau+1) + 2(n+1)^2 block products"},
            "path": "batched dyson_timestep with shared Sigma, size1 8: Matsubara integral and lesser "
                    "convolution as complex GEMMs on the FP64 matrix cores"}


def run_multiorbital_full(args):
    """The whole 16x16-grid workload on one GPU: 256 k-points, norb 8, beta 20, ntau 600,
    nt 1000, h 0.01, Sigma = 0.09 G_loc, self-consistent to 1e-12 at every timestep.  The
    k-points are herm_matrix_window objects: whole G^tv histories (158 GB) and the last 16
    retarded rows and lesser columns on the device; k-point 0 streams its rows to pinned
    host memory, the others drop them (G_loc holds every observable)."""
    import torch
    from nessi_amd import multiorbital as mo
    nkx, nt, W = 16, args.mo_full_nt, 16
    log("16x16 workload: 256 k-points, norb 8, ntau 600, nt %d, window %d" % (nt, W))
    T0 = time.perf_counter()
    L = mo.MultiOrbitalLattice(nkx, 8, 20.0, 600, nt, 0.01, g2=0.09, window=W, host_k=(0,))
    torch.cuda.synchronize()
    log("16x16 workload: allocated %.1f GB on the device" % (torch.cuda.memory_allocated() / 1e9))
    t0 = time.perf_counter()
    mat_it = L.matsubara(1e-12, 50, progress=lambda it: log("16x16 workload: Matsubara iteration %d" % it))
    torch.cuda.synchronize()
    t_mat = time.perf_counter() - t0
    t0 = time.perf_counter()
    start_it = L.start(1e-12, 50)
    torch.cuda.synchronize()
    t_start = time.perf_counter() - t0
    log("16x16 workload: Matsubara %.1f s (%d iterations), start %.1f s" % (t_mat, mat_it, t_start))
    t0 = time.perf_counter()
    iters, flops = 0, 0.0
    for n in range(L.k + 1, nt + 1):
        it = L.step(n, 1e-12, 50)
        iters += it
        e_n = (L.ntau + 1.0) ** 2 + 2.0 * (n + 1) * (L.ntau + 1) + 2.0 * (n + 1) ** 2
        flops += it * L.nk * e_n * 8.0 * 8 ** 3
        if n % 25 == 0:
            torch.cuda.synchronize()
            log("16x16 workload: timestep %d, %.1f s" % (n, time.perf_counter() - t0))
    L.finish()
    torch.cuda.synchronize()
    secs = time.perf_counter() - t0
    steps = nt - L.k
    rho = L.Gloc.density_matrix(nt)
    return {"workload": "16x16 k-grid, norb 8, beta 20, ntau 600, nt %d, h 0.01, Sigma = 0.09 G_loc, "
                        "H(k,t) = H_rand + eps(k - A(t))" % nt,
            "value": L.nk * steps / secs, "unit": "k-timesteps/s", "timed_timesteps": [L.k + 1, nt],
            "corrector_iterations": iters, "seconds": secs, "matsubara_seconds": t_mat,
            "matsubara_iterations": mat_it, "start_seconds": t_start, "start_iterations": start_it,
            "wall_seconds": time.perf_counter() - T0,
            "device_GB_peak": torch.cuda.max_memory_allocated() / 1e9,
            "window_timesteps": W, "streamed_to_host": "k-point 0 (ret, les)",
            "G_loc_density_trace_at_nt": float(np.trace(rho).real),
            "fp64": {"algorithmic_flops": flops, "achieved_TFLOPs": flops / secs / 1e12,
                     "frac": flops / secs / 1e12 / PEAK_FP64_TFLOPS}}


def run_serving(args):
    """Poisson arrivals of requests (the uneven-batch distribution with nt up to 2000) into
    the batch server (nessi_amd/batch_server.py) with a 200 GB budget.  Per arrival rate:
    completed solves/s and the latency from the scheduled arrival to the tick that saw the
    request's last timestep complete."""
    import torch
    from nessi_amd import batch_server, workloads
    rates = [float(x) for x in args.serving_rates.split(",")]
    n = args.serving_requests
    out = []
    for lam in rates:
        # the same requests in the same order at every rate, arriving on one Poisson
        # sequence scaled by 1/rate, so that the latencies differ by the load alone
        rs = workloads.uneven_batch(n, seed=100, nt_hi=args.serving_nt_max)
        arr = workloads.poisson_arrivals(lam, n, seed=200)
        S = batch_server.BatchServer(200 << 30)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sched, lat = {}, {}
        i, live_sum, ticks = 0, 0, 0
        last_log = 0.0
        while len(lat) < n:
            if time.perf_counter() - t0 - last_log > 10.0:
                last_log = time.perf_counter() - t0
                log("serving rate %g/s: %d submitted, %d done" % (lam, i, len(lat)))
            now = time.perf_counter() - t0
            while i < n and arr[i] <= now:
                sched[S.submit(rs[i])] = (arr[i], now)
                i += 1
            live, queued = S.step()
            if live or queued:
                live_sum += live
                ticks += 1
            elif i < n:
                time.sleep(max(0.0, min(arr[i] - (time.perf_counter() - t0), 0.002)))
            for rid, srv_lat in S.poll():
                a, sub = sched[rid]
                lat[rid] = srv_lat + (sub - a)   # the server counts from submission
                S.release(rid)
        span = time.perf_counter() - t0
        log("serving rate %g/s: %d requests in %.1f s" % (lam, n, span))
        timesteps = S.timesteps
        S.close()
        v = np.array(sorted(lat.values()))
        out.append({"rate_per_s": lam, "requests": n, "nt_max": args.serving_nt_max,
                    "completed_per_s": n / span, "seconds": span,
                    "p50_latency_s": float(np.percentile(v, 50)),
                    "p90_latency_s": float(np.percentile(v, 90)),
                    "p99_latency_s": float(np.percentile(v, 99)),
                    "mean_live": live_sum / max(ticks, 1), "timesteps": timesteps,
                    "solve_timesteps_per_s": timesteps / span,
                    "path": "C++ batch server (cntrx_server_*): one batched dyson_timestep call over "
                            "every live solve per tick"})
    return out


def gloc_slices(g, t0, t1):
    """G_loc's timesteps t0..t1 (ret rows, les columns, tv rows) as one vector"""
    parts = []
    for t in range(t0, t1 + 1):
        o = t * (t + 1) // 2
        parts += [g["ret"][o:o + t + 1].ravel(), g["les"][o:o + t + 1].ravel(),
                  g["tv"][t * (NTAU + 1):(t + 1) * (NTAU + 1)].ravel()]
    return np.concatenate(parts)


def run_lattice(args):
    """The headline: `warmup` untimed sweeps, then `steps` timed sweeps of timesteps
    6..400 through the device-driven run (cntrx_lattice_run), with live HIP-event timing
    of the four batched Dyson kernels; then the plain per-k loop over the same sweep, the
    host oracle on a bounded prefix and the parity check."""
    import torch
    from nessi_amd import lattice as latmod
    epsk = latmod.chain_1d(NK)
    W, KS = args.warmup, args.steps
    last = NT
    per_sweep = NK * (last - FIRST + 1)

    log("lattice: setup (equilibrium, bootstrap)")
    A = latmod.Lattice(epsk, U, BETA, NTAU, NT, H, K)
    A.equilibrium()
    boot_it = A.bootstrap(TOL, MAXITER)
    torch.cuda.synchronize()
    its_ref, gl_ref = None, None
    for w in range(W):
        its, _ = A.run(FIRST, last, TOL, MAXITER)
        if its_ref is None:
            its_ref = its.copy()
            gl_ref = A.fetch(-1)
    torch.cuda.synchronize()
    log("lattice: %d warmup sweeps done" % W)
    # live timing inside the timed sweeps: every fourth chunk (one timestep's iterations)
    # brackets one of the four batched Dyson kernels of its first iteration with HIP
    # events on the lattice stream, the four in turn
    A.profile(A.PROFILE_SAMPLED)
    torch.cuda.synchronize()
    sweep_iters = []
    T0 = time.perf_counter()
    for s in range(KS):
        its, _ = A.run(FIRST, last, TOL, MAXITER)
        sweep_iters.append(its)
    torch.cuda.synchronize()
    T1 = time.perf_counter()
    elapsed = T1 - T0
    log("lattice: %d timed sweeps of timesteps %d..%d in %.2f s" % (KS, FIRST, last, elapsed))
    kq, kt, kms = A.kernel_samples()
    A.profile(A.PROFILE_OFF)
    if its_ref is None:
        its_ref = sweep_iters[0]
    gl_end = A.fetch(-1)
    same_iters = all(np.array_equal(x, its_ref) for x in sweep_iters)
    same_gloc = True if gl_ref is None else bool(np.array_equal(gloc_slices(gl_ref, 0, last),
                                                                gloc_slices(gl_end, 0, last)))
    iters_per_sweep = int(np.sum(its_ref))
    value = per_sweep * KS / elapsed

    # roofline of the kernel with the largest average launch time on the iteration's path.
    # Every kernel runs once per corrector iteration.  In the fused lattice the mixed
    # component and the retarded row are one launch (k_tvret1, timing slot 0; slot 1 times
    # nothing); otherwise k_retrec runs on a second stream beside k_tv and the pair costs
    # the iteration the later of the two to end.
    fused = A.fused()
    if fused:
        names = ("k_tvret1", None, "k_lesconv", "k_lesrec")
        on_path = ("k_tvret1", "k_lesconv", "k_lesrec")
    else:
        names = ("k_tv", "k_retrec", "k_lesconv", "k_lesrec")
        on_path = ("k_tv", "k_lesconv", "k_lesrec")
    per = {}
    for q, name in enumerate(names):
        if name is None:
            continue
        sel = kq == q
        n_l = int(sel.sum())
        ms_sum = float(kms[sel].sum())
        b_sum = sum(kernel_bytes(int(t), NTAU, NK)[name] for t in kt[sel])
        # mean over every launch of a sweep (one per corrector iteration): the window
        # a counter run over whole sweeps averages over
        b_sweep = sum(int(n) * kernel_bytes(FIRST + i, NTAU, NK)[name] for i, n in enumerate(its_ref))
        per[name] = {"launches_timed": n_l, "avg_launch_ms": ms_sum / max(n_l, 1),
                     "algorithmic_bytes_per_launch": b_sum / max(n_l, 1),
                     "algorithmic_bytes_per_launch_sweep_mean": b_sweep / max(int(np.sum(its_ref)), 1),
                     "GBs": (b_sum / (ms_sum * 1e-3) / 1e9) if ms_sum > 0 else 0.0}
    dom = max(on_path, key=lambda n: per[n]["avg_launch_ms"])
    traffic, traffic_window = pmc_traffic(dom)
    achieved = per[dom]["GBs"]
    roofline = {"kernel": dom, "bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS,
                "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS, "traffic": traffic,
                "traffic_source": "profiles/lattice_nk128_nt400_pmc.md (%s)" % traffic_window,
                "avg_launch_ms": per[dom]["avg_launch_ms"], "launches": per[dom]["launches_timed"],
                "algorithmic_bytes_per_launch": per[dom]["algorithmic_bytes_per_launch"],
                "timing": "HIP events on the lattice stream around one batched kernel in the first "
                          "corrector iteration of every fourth chunk (one timestep's iterations), "
                          "the kernels in turn, over every timed sweep (timesteps %d..%d)"
                          % (FIRST, last),
                "selection": ("largest average launch time among %s; " % ", ".join(on_path)) +
                             ("k_tvret1 is the mixed component (G^tv history stream) and the retarded "
                              "row in one launch, its bytes the sum of the two" if fused else
                              "k_retrec overlaps k_tv on a second stream and is latency-bound, "
                              "so the pair is represented by k_tv"),
                "kernels": per}

    # ---- plain per-k loop over the same sweep -----------------------------------------
    plain = None
    if not args.no_plain:
        plast = min(args.plain_last, last)
        B = latmod.Lattice(epsk, U, BETA, NTAU, NT, H, K)
        B.equilibrium()
        B.bootstrap(TOL, MAXITER)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for t in range(FIRST, plast + 1):
            B.step(t, False, TOL, MAXITER)
        torch.cuda.synchronize()
        plain_time = time.perf_counter() - t0
        gb = B.fetch(-1)
        B.close()
        if plast == last:
            batched_same = elapsed / KS
        else:
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            A.run(FIRST, plast, TOL, MAXITER)
            torch.cuda.synchronize()
            batched_same = time.perf_counter() - t0
        identical = bool(np.array_equal(gloc_slices(gl_end, 0, plast), gloc_slices(gb, 0, plast)))
        n_k = NK * (plast - FIRST + 1)
        plain = {"timesteps": [FIRST, plast], "plain_value": n_k / plain_time,
                 "batched_value_same_steps": n_k / batched_same, "speedup": plain_time / batched_same,
                 "plain_seconds": plain_time, "bitwise_identical_G_loc": identical,
                 "plain_path": "cntrx_lattice_step(batched=0): per iteration, the local Sigma, then "
                               "cntrx_dyson_timestep once per k, then G_loc and the distance"}
        log("lattice: plain per-k loop over timesteps %d..%d in %.1f s" % (FIRST, plast, plain_time))

    # ---- CPU baseline + parity on a bounded prefix ------------------------------------
    cpu, parity = None, None
    if not args.no_cpu:
        from oracle import oracle
        info = host_info()
        nstop = min(K + args.cpu_steps, last)
        thr = max(1, min(args.cpu_threads, info["affinity_cpus"] or 1))
        kchk = (0, 37)
        dens, it_cpu, secs, gl, gks = oracle.lattice_run(epsk, U, BETA, NTAU, NT, H, K, nstop, TOL,
                                                         MAXITER, thr, True, kidx=kchk)
        log("lattice: host oracle, %d threads, timesteps %d..%d: %.1f s" % (thr, FIRST, nstop, secs))
        n1 = min(K + args.cpu1_steps, last)
        _, _, secs1, _ = oracle.lattice_run(epsk, U, BETA, NTAU, NT, H, K, n1, TOL, MAXITER, 1, False,
                                            setup_threads=thr)
        log("lattice: host oracle, 1 thread, timesteps %d..%d: %.1f s" % (FIRST, n1, secs1))
        cpu = {"value": NK * (nstop - K) / secs, "unit": "k-timesteps/s", "cores": thr,
               "kind": "port",
               "sample": "same workload, timesteps %d..%d of %d, all %d k-points, OpenMP over k "
                         "(setup excluded), %.1f s" % (FIRST, nstop, NT, NK, secs),
               "single_core": {"value": NK * (n1 - K) / secs1, "unit": "k-timesteps/s",
                               "sample": "timesteps %d..%d, all %d k-points, 1 thread, %.1f s"
                                         % (FIRST, n1, NK, secs1)},
               "host": info}
        errs = {"G_loc": herm_rel_err(gl_end, gl, nstop, NTAU, ("ret", "les", "tv"))}
        for q, X in zip(kchk, gks):
            errs["G_k[%d]" % q] = herm_rel_err(A.fetch(q), X, nstop, NTAU)
        parity = {"max_rel_err_vs_oracle": errs,
                  "timesteps_checked": "0..%d of the last timed sweep: G_loc (ret, les, tv; the "
                                       "lattice's Sigma^M is 0, so G_loc^M is not formed) and G_k "
                                       "for k index %s (ret, les, tv, mat)" % (nstop, list(kchk)),
                  "tolerance": 1e-12,
                  "pass": all(e <= 1e-12 for d in errs.values() for e in d.values())}
    A.close()
    return {
        "value": value, "elapsed": elapsed, "KS": KS, "W": W, "per_sweep": per_sweep,
        "iters_per_sweep": iters_per_sweep, "boot_it": boot_it, "roofline": roofline,
        "plain": plain, "cpu": cpu, "parity": parity,
        "sweeps": {"k_timesteps_per_sweep": per_sweep, "timesteps_per_sweep": [FIRST, last],
                   "corrector_iterations_per_sweep": iters_per_sweep,
                   "iteration_counts_identical_every_sweep": bool(same_iters),
                   "G_loc_bitwise_identical_to_first_sweep": same_gloc,
                   "ms_per_timestep": 1e3 * elapsed / (KS * (last - FIRST + 1))},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed sweeps (timesteps 6..400 each)")
    ap.add_argument("--warmup", type=int, default=2, help="untimed sweeps before the timed ones")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--plain-last", type=int, default=NT,
                    help="last timestep of the plain per-k loop's sweep")
    ap.add_argument("--cpu-steps", type=int, default=100)
    ap.add_argument("--cpu1-steps", type=int, default=15)
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-uneven", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="run the uneven batch with the autotuner off")
    ap.add_argument("--uneven-parity-steps", type=int, default=40)
    ap.add_argument("--no-multiorbital", action="store_true")
    ap.add_argument("--mo-nkx", type=int, default=4)
    ap.add_argument("--mo-steps", type=int, default=30)
    ap.add_argument("--mo-full", action="store_true",
                    help="also run the whole 16x16-grid workload (256 k-points, nt 1000; "
                         "several minutes)")
    ap.add_argument("--mo-full-nt", type=int, default=1000)
    ap.add_argument("--only-mo-full", action="store_true",
                    help="run only the whole 16x16-grid workload and print its record")
    ap.add_argument("--only-serving", action="store_true",
                    help="run only the serving sweep (--serving-rates, --serving-requests)")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="after the lattice leg, one device-to-device copy of this many bytes: "
                         "the calibration kernel for HBM byte counters (rocprofv3 --pmc)")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--serving-nt-max", type=int, default=2000)
    ap.add_argument("--plain-solves", type=int, default=8)
    ap.add_argument("--cpu-uneven-steps", type=int, default=80)
    ap.add_argument("--serving-rates", default="1,4")
    ap.add_argument("--serving-requests", type=int, default=12)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the k-batched lattice runs on a single device; "
              "--gpus %d is not supported (multi-GPU sharding is design-only)." % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    if args.steps < 1 or args.warmup < 0:
        print("--steps must be >= 1 and --warmup >= 0", file=sys.stderr)
        sys.exit(2)
    import torch

    if args.only_mo_full:
        print(json.dumps({"multiorbital_16x16_full": run_multiorbital_full(args)}))
        return
    if args.only_serving:
        print(json.dumps({"serving": run_serving(args)}))
        return
    lat = run_lattice(args)
    if args.copy_probe > 0:
        src = torch.ones(args.copy_probe, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log("copy probe: %d bytes" % args.copy_probe)
        del src, dst
    uneven = None if args.no_uneven else run_uneven(args)
    multi = None if args.no_multiorbital else run_multiorbital(args)
    multi_full = run_multiorbital_full(args) if args.mo_full else None
    serving = None if args.no_serving else run_serving(args)
    p50 = p99 = None
    if serving:
        ref = next((x for x in serving if x["rate_per_s"] == 1.0), serving[0])
        p50, p99 = ref["p50_latency_s"], ref["p99_latency_s"]

    out = {
        "metric": "k·Dyson timesteps/s batched vs per-k plain at ≤1e-12 err; p50/p99 solve latency",
        "value": lat["value"],
        "unit": "k-timesteps/s",
        "n_gpus": 1,
        "steps": lat["KS"],
        "warmup": lat["W"],
        "ms_per_step": 1e3 * lat["elapsed"] / lat["KS"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128",
        "data": "synthetic (1D chain, seeded by construction)",
        "config": {"workload": "k-batched 1D tight-binding chain, 2nd-order local self-energy, U 0->2 quench",
                   "step": "one full propagation sweep: timesteps %d..%d of all %d k-points, "
                           "predictor and every corrector iteration" % (FIRST, NT, NK),
                   "nk": NK, "norb": NORB, "beta": BETA, "ntau": NTAU, "nt": NT, "h": H,
                   "SolveOrder": K, "tol": TOL, "timed_timesteps": [FIRST, NT]},
        "sweeps": lat["sweeps"],
        "dyson_calls_per_s": NK * lat["iters_per_sweep"] * lat["KS"] / lat["elapsed"],
        "bootstrap_iterations": lat["boot_it"],
        "batched_vs_plain": lat["plain"],
        "roofline": lat["roofline"],
        "cpu_baseline": lat["cpu"],
        "parity": lat["parity"],
        "p50_latency_s": p50,
        "p99_latency_s": p99,
        "uneven_batch": uneven,
        "multiorbital_k_batch": multi,
        "multiorbital_16x16_full": multi_full,
        "serving": serving,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
