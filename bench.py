"""Benchmark: k-batched Kadanoff-Baym time stepping on one MI355X.

Workload (the reference's k-batched lattice, config 2 of the measurement plan):
1D tight-binding chain eps_k = -2 cos(k), Nk = 128, norb = 1, beta = 10, ntau = 400,
nt = 400, h = 0.025, SolveOrder 5, local second-order self-energy
Sigma = U^2 G_loc(t,t')^2 G_loc(t',t) from Bubble1/Bubble2 with an interaction quench
U: 0 -> 2 at t = 0, self-consistent to distance_norm2 < 1e-12 at every timestep.

A "step" is one physical timestep of all Nk k-points (predictor + corrector iterations).
`value` = Nk * timed steps / wall time of the timed region (batched launches, all data
resident in HBM).  The same steps run through the plain per-k entry point
(cntrx_dyson_timestep once per k) give the batched-vs-plain ratio, and the CPU oracle
(OpenMP over k, the restated reference algorithm) runs a bounded prefix of the same
workload on the host cores for the CPU baseline and the parity check.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

NK, NORB, BETA, NTAU, NT, H, K, U = 128, 1, 10.0, 400, 400, 0.025, 5, 2.0
TOL, MAXITER = 1e-12, 50
PEAK_HBM_GBS = 8000.0


def kernel_bytes(t, ntau, nk):
    """algorithmic HBM bytes of one batched launch of each kernel at timestep t
    (complex128 = 16 B; shared Sigma counted once per launch)."""
    row = 16.0 * (ntau + 1)
    tri = 16.0 * (t + 1) * (t + 2) / 2
    col = 16.0 * (t + 1)
    return {
        "k_tv": nk * (t + 2) * row + row + col,          # G^tv history + G^M + out; Sigma rows
        "k_retrec": tri + nk * 2 * col,                    # Sigma^R triangle; H, out row
        "k_lesconv": (t + 1) * row + tri + nk * (row + 2 * col),  # Sigma^tv hist, Sigma^< tri
        "k_lesrec": tri + nk * 2 * col,                    # Sigma^R triangle; gles, out col
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--plain-steps", type=int, default=12)
    ap.add_argument("--cpu-steps", type=int, default=40)
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the k-batched lattice runs on a single device; "
              "--gpus %d is not supported (multi-GPU sharding is design-only)." % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    from nessi_amd import lattice as latmod

    W = args.warmup
    first = K + 1
    K_steps = args.steps if args.steps is not None else NT - K - W
    K_steps = max(1, min(K_steps, NT - K - W))
    epsk = latmod.chain_1d(NK)

    # ---- batched run over the whole propagation ----------------------------------
    A = latmod.Lattice(epsk, U, BETA, NTAU, NT, H, K)
    A.equilibrium()
    boot_it = A.bootstrap(TOL, MAXITER)
    torch.cuda.synchronize()
    step_time, step_iters = {}, {}
    for t in range(first, first + W):
        t0 = time.perf_counter()
        it, _ = A.step(t, True, TOL, MAXITER)
        torch.cuda.synchronize()
        step_time[t], step_iters[t] = time.perf_counter() - t0, it
    # live timing over the timed region: every 8th corrector iteration brackets one of
    # the four batched Dyson kernels with HIP events (the four in turn) and replays the
    # rest of that iteration as graphs; the other iterations run as the full graph
    A.profile(A.PROFILE_SAMPLED)
    torch.cuda.synchronize()
    T0 = time.perf_counter()
    for t in range(first + W, first + W + K_steps):
        t0 = time.perf_counter()
        it, _ = A.step(t, True, TOL, MAXITER)
        torch.cuda.synchronize()
        step_time[t], step_iters[t] = time.perf_counter() - t0, it
    torch.cuda.synchronize()
    T1 = time.perf_counter()
    kq, kt, kms = A.kernel_samples()
    elapsed = T1 - T0
    timed = range(first + W, first + W + K_steps)
    iters_timed = sum(step_iters[t] for t in timed)
    value = NK * K_steps / elapsed

    # roofline of the kernel with the largest average launch time (every kernel runs
    # once per corrector iteration, so that is the largest share of the step)
    names = ("k_tv", "k_retrec", "k_lesconv", "k_lesrec")
    per = {}
    for q, name in enumerate(names):
        sel = kq == q
        n_l = int(sel.sum())
        ms_sum = float(kms[sel].sum())
        b_sum = sum(kernel_bytes(int(t), NTAU, NK)[name] for t in kt[sel])
        per[name] = {"launches_timed": n_l, "avg_launch_ms": ms_sum / max(n_l, 1),
                     "algorithmic_bytes_per_launch": b_sum / max(n_l, 1),
                     "GBs": (b_sum / (ms_sum * 1e-3) / 1e9) if ms_sum > 0 else 0.0}
    dom = max(per, key=lambda n: per[n]["avg_launch_ms"])
    achieved = per[dom]["GBs"]
    roofline = {"kernel": dom, "bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS,
                "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS, "traffic": None,
                "avg_launch_ms": per[dom]["avg_launch_ms"], "launches": per[dom]["launches_timed"],
                "algorithmic_bytes_per_launch": per[dom]["algorithmic_bytes_per_launch"],
                "timing": "HIP events on the lattice stream around one batched kernel in every "
                          "8th corrector iteration, the four kernels in turn, whole timed region",
                "kernels": per}

    # ---- plain per-k loop over the first steps --------------------------------------
    P = max(1, min(args.plain_steps, W + K_steps))
    B = latmod.Lattice(epsk, U, BETA, NTAU, NT, H, K)
    B.equilibrium()
    B.bootstrap(TOL, MAXITER)
    torch.cuda.synchronize()
    plain_time = 0.0
    for t in range(first, first + P):
        t0 = time.perf_counter()
        B.step(t, False, TOL, MAXITER)
        torch.cuda.synchronize()
        plain_time += time.perf_counter() - t0
    batched_same = sum(step_time[t] for t in range(first, first + P))
    ga, gb = A.fetch(-1), B.fetch(-1)
    last = first + P - 1

    def sl(g, t):
        o = t * (t + 1) // 2
        return np.concatenate([g["ret"][o:o + t + 1].ravel(), g["les"][o:o + t + 1].ravel(),
                               g["tv"][t * (NTAU + 1):(t + 1) * (NTAU + 1)].ravel()])

    identical = all(np.array_equal(sl(ga, t), sl(gb, t)) for t in range(0, last + 1))
    B.close()
    plain = {"steps": P, "plain_value": NK * P / plain_time, "batched_value_same_steps": NK * P / batched_same,
             "speedup": plain_time / batched_same, "bitwise_identical_G_loc": bool(identical)}

    # ---- CPU baseline + parity on a bounded prefix ----------------------------------
    cpu, parity = None, None
    if not args.no_cpu:
        from oracle import oracle
        nstop = K + args.cpu_steps
        thr = max(1, min(args.cpu_threads, os.cpu_count() or 1))
        dens, it_cpu, secs, gl = oracle.lattice_run(epsk, U, BETA, NTAU, NT, H, K, nstop, TOL,
                                                    MAXITER, thr, True)
        cpu = {"value": NK * (nstop - K) / secs, "unit": "k-timesteps/s", "cores": thr,
               "kind": "port",
               "sample": "same workload, timesteps %d..%d of %d, all %d k-points, OpenMP over k "
                         "(setup excluded)" % (K + 1, nstop, NT, NK)}
        errs = {}
        for name, get in (("ret", lambda g, t: g["ret"][t * (t + 1) // 2:t * (t + 1) // 2 + t + 1]),
                          ("les", lambda g, t: g["les"][t * (t + 1) // 2:t * (t + 1) // 2 + t + 1]),
                          ("tv", lambda g, t: g["tv"][t * (NTAU + 1):(t + 1) * (NTAU + 1)])):
            x = np.concatenate([get(ga, t).ravel() for t in range(0, nstop + 1)])
            gh = {"ret": gl.ret, "les": gl.les, "tv": gl.tv}
            y = np.concatenate([get(gh, t).ravel() for t in range(0, nstop + 1)])
            errs[name] = float(np.abs(x - y).max() / np.abs(y).max())
        parity = {"max_rel_err_vs_oracle": errs, "timesteps_checked": "0..%d (G_loc, all components)" % nstop,
                  "tolerance": 1e-12, "pass": all(e <= 1e-12 for e in errs.values())}
    A.close()

    out = {
        "metric": "k·Dyson timesteps/s batched vs per-k plain at <=1e-12 err",
        "value": value,
        "unit": "k-timesteps/s",
        "n_gpus": 1,
        "steps": K_steps,
        "warmup": W,
        "ms_per_step": 1e3 * elapsed / K_steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128",
        "data": "synthetic (1D chain, seeded by construction)",
        "config": {"workload": "k-batched 1D tight-binding chain, 2nd-order local self-energy, U 0->2 quench",
                   "nk": NK, "norb": NORB, "beta": BETA, "ntau": NTAU, "nt": NT, "h": H,
                   "SolveOrder": K, "tol": TOL, "timed_timesteps": [first + W, first + W + K_steps - 1]},
        "corrector_iterations_timed": iters_timed,
        "dyson_calls_per_s": NK * iters_timed / elapsed,
        "bootstrap_iterations": boot_it,
        "batched_vs_plain": plain,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "parity": parity,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
