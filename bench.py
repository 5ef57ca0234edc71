"""Benchmark: BEAGLE likelihood evaluations on one MI355X.

Workload (the configuration the headline metric is quoted on): 4-state
GTR + Gamma(4 categories, shape 0.5), 128 taxa, 100,000 unique site
patterns of uniform random tip states, pattern weights 1 + Poisson(2),
branch lengths ~ Exp(mean 0.1), two tree shapes: a 127-deep caterpillar
and a balanced tree.  Synthetic data, seed 1.

One step = one full FP64 log-likelihood evaluation of each tree through
the beagle.h API: transition matrices for all 254 branches from the
eigen-system, the 127-operation post-order traversal, the root
log-likelihood (FP64 sum read back to the host).  value = evaluations per
second (2 per step).  Inputs are resident in HBM before timing starts.

Also reported: the FP32 path with exact power-of-two rescaling (its
relative logL error vs FP64), a logL + full gradient evaluation (pre-order
pass plus first and second derivatives for all 254 branches), the fused
post-order kernel's roofline (algorithmic bytes per launch over its
duration measured with HIP events on the library's stream), and the
reference's threaded CPU plugin timed on this host.
"""
import argparse
import json
import os
import subprocess
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "beagle-lib_amd"))

HBM_PEAK_GBS = 8000.0
N_TAXA, N_PATTERNS, SEED = 128, 100_000, 1
METRIC = "logL+gradient evals/s & partials GB/s vs plain FP64, at ≤1e-6 rel err"


def postorder_bytes(prob, w):
    """Algorithmic HBM bytes of one post-order traversal: per operation the
    destination write C*P*S*w, each child (4*P for tip states, C*P*S*w for
    partials) and the two C*S*S*w matrices."""
    C, P, S = prob.C, prob.P, prob.S
    n = prob.tree.n
    total = 0
    for v in prob.tree.postorder():
        a, b = prob.tree.children[v]
        total += C * P * S * w + 2 * C * S * S * w
        total += sum(4 * P if x < n else C * P * S * w for x in (a, b))
    return total


def problems():
    from beagle_amd import synth
    return [synth.make_problem(SEED, N_TAXA, N_PATTERNS, S=4, shape=s)
            for s in ("caterpillar", "balanced")]


def cpu_baseline(budget_s):
    """Times the reference CPU plugin (threaded SSE, double) on this host's
    cores, on the same workload, for about budget_s seconds of evaluations;
    falls back to the oracle (single thread) if the reference build is not
    present.  Runs in its own process (both libraries export beagle*)."""
    import numpy as np
    from beagle_amd import api, driver
    ref_path = os.path.join(REPO, "oracle", "_ref", "libhmsbeagle.so")
    cores = min(len(os.sched_getaffinity(0)), 16)
    prob = problems()[1]
    if os.path.exists(ref_path):
        lib = api.Library(ref_path, "beagle")
        flags = api.FLAG_PROCESSOR_CPU | api.FLAG_VECTOR_SSE | api.FLAG_THREADING_CPP
        ev = driver.Evaluator(lib, prob, preorder=False, extra_flags=flags, transpose="none")
        lib.call("SetCPUThreadCount", ev.inst.h, cores)
        kind, name = "reference", ev.inst.impl_name
    else:
        lib = api.Library(os.path.join(REPO, "oracle", "liboracle.so"), "orc")
        ev = driver.Evaluator(lib, prob, preorder=False, transpose="none")
        kind, name, cores = "port", "oracle", 1
    logl = ev.log_likelihood()
    t0 = time.perf_counter()
    n = 0
    while True:
        ev.log_likelihood()
        n += 1
        if time.perf_counter() - t0 > budget_s and n >= 3:
            break
    dt = time.perf_counter() - t0
    ev.close()
    return {"value": n / dt, "unit": "evals/s", "cores": cores, "kind": kind,
            "sample": f"{n} FP64 logL evaluations of the balanced 128-taxon, 100,000-pattern "
                      f"tree with {name} ({dt:.1f} s)", "logL": logl}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-baseline-only", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a plain 1 GiB -> 1 GiB elementwise copy kernel, the "
                         "reference size for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE")
    args = ap.parse_args()

    if args.cpu_baseline_only:
        print(json.dumps(cpu_baseline(args.cpu_budget)))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(patterns shard across GPUs with one all-reduce per evaluation; "
              "see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import beagle_amd
    from beagle_amd import api, driver

    assert torch.cuda.is_available(), "needs an MI355X"
    lib = beagle_amd.load()
    probs = problems()
    async_flag = api.FLAG_COMPUTATION_ASYNCH

    # ---- FP64 plain path (the timed metric) ----
    evs = [driver.Evaluator(lib, p, preorder=False, extra_flags=async_flag) for p in probs]
    streams = [torch.cuda.ExternalStream(ev.inst.stream()) for ev in evs]
    for _ in range(args.warmup):
        for ev in evs:
            ev.log_likelihood()
    logl64 = [ev.log_likelihood() for ev in evs]
    ev_pairs = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        for ev, st in zip(evs, streams):
            ev.update_matrices()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(st)
            ev.post_order()
            e1.record(st)
            ev.root()
            ev_pairs.append((e0, e1))
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms_per_step = dt * 1e3 / args.steps
    evals_per_s = 2 * args.steps / dt
    kernel_ms = float(np.mean([a.elapsed_time(b) for a, b in ev_pairs]))
    bytes64 = postorder_bytes(probs[0], 8)
    achieved = bytes64 / (kernel_ms * 1e-3) / 1e9
    for ev in evs:
        ev.close()

    # ---- FP32 with exact power-of-two rescaling ----
    evs = [driver.Evaluator(lib, p, preorder=False, precision="single", scaling="auto",
                            extra_flags=async_flag) for p in probs]
    for _ in range(args.warmup):
        for ev in evs:
            ev.log_likelihood()
    logl32 = [ev.log_likelihood() for ev in evs]
    n32 = max(args.steps // 2, 5)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(n32):
        for ev in evs:
            ev.log_likelihood()
    torch.cuda.synchronize()
    fp32_eps = 2 * n32 / (time.perf_counter() - t0)
    for ev in evs:
        ev.close()
    err32 = max(abs(a - b) / abs(b) for a, b in zip(logl32, logl64))

    # ---- logL + full gradient (FP64): API call sequence, and the fused
    # pre-order + derivative extension ----
    ev = driver.Evaluator(lib, probs[1], extra_flags=async_flag)
    ng = max(args.steps // 5, 3)
    rates = {}
    for name, fn in (("api", ev.gradient), ("fused", ev.fused_gradient)):
        for _ in range(2):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(ng):
            res = fn()
        torch.cuda.synchronize()
        rates[name] = (ng / (time.perf_counter() - t0), res)
    g_api, g_fused = rates["api"][1], rates["fused"][1]
    grad_err = max(float(np.max(np.abs(g_api[1] - g_fused[1]) / np.maximum(np.abs(g_api[1]), 1e-12))),
                   float(np.max(np.abs(g_api[2] - g_fused[2]) / np.maximum(np.abs(g_api[2]), 1e-12))))
    ev.close()
    # FP32 auto-scaled fused gradient vs the FP64 one
    ev = driver.Evaluator(lib, probs[1], extra_flags=async_flag, precision="single",
                          scaling="auto")
    for _ in range(2):
        ev.fused_gradient()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(ng):
        g32 = ev.fused_gradient()
    torch.cuda.synchronize()
    grad32_eps = ng / (time.perf_counter() - t0)
    ev.close()
    grad32_err = float(np.max(np.abs(g32[1] - g_fused[1])) / np.max(np.abs(g_fused[1])))

    if args.calibrate_copy:
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")  # 1 GiB
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        del src, dst

    cpu = None
    if not args.no_cpu_baseline:
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-baseline-only",
                            "--cpu-budget", str(args.cpu_budget)], capture_output=True, text=True,
                           timeout=600)
        if r.returncode == 0:
            cpu = json.loads(r.stdout.strip().splitlines()[-1])
            cpu["logL_rel_err_vs_gpu"] = abs(cpu.pop("logL") - logl64[1]) / abs(logl64[1])
        else:
            print(r.stderr[-2000:], file=sys.stderr)

    traffic = None
    tpath = os.path.join(REPO, "profiles", "post4_fp64_traffic.json")
    if os.path.exists(tpath):
        traffic = json.load(open(tpath)).get("hbm_bytes_per_launch")

    out = {
        "metric": METRIC,
        "value": evals_per_s,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generators; no dataset)",
        "config": {"workload": "4-state GTR+G4, 128 taxa, 100,000 patterns, caterpillar + "
                               "balanced trees, FP64 post-order + root logL",
                   "taxa": N_TAXA, "patterns": N_PATTERNS, "categories": 4,
                   "evals_per_step": 2},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_post4<double> (fused 127-op post-order traversal)",
                     "kernel_ms": kernel_ms, "algorithmic_bytes": bytes64},
        "partials_GBps_fp64": achieved,
        "fp32_auto": {"evals_per_s": fp32_eps, "speedup_vs_fp64": fp32_eps / evals_per_s,
                      "logL_rel_err_vs_fp64": err32},
        "logL_plus_gradient": {
            "what": "post-order + root + pre-order + dlogL/dt and d2logL/dt2 for all 254 "
                    "branches, balanced tree",
            "fp64_api_evals_per_s": rates["api"][0],
            "fp64_fused_evals_per_s": rates["fused"][0],
            "fused_vs_api_max_rel_diff": grad_err,
            "fp32_auto_fused_evals_per_s": grad32_eps,
            "fp32_gradient_max_err_rel_to_max": grad32_err,
            "speedup_fp32_fused_vs_fp64_api": grad32_eps / rates["api"][0]},
        "logL": logl64,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
