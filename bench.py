#!/usr/bin/env python3
"""Propagation benchmark on one MI355X.

Workload: the single large sparse MIP (m = 1,000,000 rows, n = 500,000
columns, about 20M nonzeros, power-law row lengths up to 50,000; see
gdp_amd.instances.powerlaw_mip), propagated to its fixpoint by the persistent
megakernel.  A step is one propagation of the root bounds to the fixpoint:
restore the root bounds on the device, run all rounds.  The matrix and the
bounds are resident in HBM before the timed region starts.

Prints one JSON line:
  value      propagation throughput, nnz x rounds / second over the timed steps
  roofline   HBM roofline of the megakernel: algorithmic bytes per launch
             (rounds x (12 nnz + 20 m + 20 n)) / average launch time, against
             8 TB/s; launch time measured with HIP events on the library's
             stream around each launch
  cpu_baseline  the oracle's cpu_seq on the same workload on this host
"""
import argparse
import json
import os
import statistics
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "gpu-domain-propagator_amd"))
sys.path.insert(0, os.path.join(REPO, "tests"))

HBM_PEAK_GBS = 8000.0
METRIC = "propagation nnz/s & HBM-roofline %; batched nodes/s vs plain launches; p50/p99"
SYNC_NAMES = {0: "cpu_loop", 1: "gpu_loop", 2: "megakernel"}


def algorithmic_bytes(inst, rounds):
    return rounds * (12 * inst.nnz + 20 * inst.m + 20 * inst.n)


def cpu_baseline(inst, runs):
    import oracle_binding
    times, rounds = [], 0
    for _ in range(runs):
        t0 = time.perf_counter()
        r = oracle_binding.seq(inst)
        times.append(time.perf_counter() - t0)
        rounds = r[2]
    t = statistics.median(times)
    return {"value": inst.nnz * rounds / t, "unit": "nnz/s", "cores": 1, "kind": "port",
            "sample": f"cpu_seq (oracle restatement, Gauss-Seidel with row marking) on the full workload, "
                      f"median of {runs} runs: {rounds} rounds in {t:.3f} s",
            "time_to_fixpoint_ms": t * 1e3, "rounds": rounds}


def omp_baseline(inst, threads):
    from gdp_amd import interface as I
    t0 = time.perf_counter()
    out = I.host_propagate(inst, "omp", threads=threads)
    t = time.perf_counter() - t0
    return {"threads": threads, "rounds": out.rounds, "time_to_fixpoint_ms": t * 1e3,
            "value": inst.nnz * out.rounds / t, "unit": "nnz/s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--sync", type=int, default=2, choices=[0, 1, 2])
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-runs", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the host baselines")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the propagation path runs on a single MI355X "
              "(multi-GPU sharding is a design note, not a measured path); rerun with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import gdp_amd
    from gdp_amd import interface as I
    from gdp_amd.instances import powerlaw_mip

    gdp_amd.load_library()
    print("bench: generating workload", file=sys.stderr, flush=True)
    inst = powerlaw_mip(seed=args.seed)
    print(f"bench: m={inst.m} n={inst.n} nnz={inst.nnz}", file=sys.stderr, flush=True)

    # Device buffers come from the library's own HIP runtime (a framework
    # with a bundled ROCm runtime cannot share the process with it).
    print(f"bench: device {I.device_name()}", file=sys.stderr, flush=True)
    root_lb = I.DeviceArray(inst.lb)
    root_ub = I.DeviceArray(inst.ub)
    lb = I.DeviceArray(inst.lb)
    ub = I.DeviceArray(inst.ub)
    prob = I.Problem(inst)

    def step():
        lb.copy_from(root_lb)   # synchronous device-to-device copies
        ub.copy_from(root_ub)
        return prob.propagate_device(lb.ptr, ub.ptr, args.sync)  # returns after the device finished

    for w in range(args.warmup):
        info = step()
        print(f"bench: warmup {w}: rounds={info.rounds} kernel_ms={info.kernel_ms:.3f}", file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    infos = [step() for _ in range(args.steps)]
    t1 = time.perf_counter()

    rounds = infos[-1].rounds
    assert all(i.rounds == rounds for i in infos), "round count changed between steps"
    total = t1 - t0
    ms_per_step = total / args.steps * 1e3
    value = inst.nnz * rounds * args.steps / total
    kernel_ms = statistics.mean(i.kernel_ms for i in infos)
    achieved = algorithmic_bytes(inst, rounds) / (kernel_ms * 1e-3) / 1e9
    print(f"bench: {rounds} rounds, {ms_per_step:.3f} ms/step, kernel {kernel_ms:.3f} ms, "
          f"{achieved:.1f} GB/s algorithmic", file=sys.stderr, flush=True)

    # parity spot check of the timed result against the oracle (cheap: one run)
    out_lb, out_ub = lb.numpy(), ub.numpy()

    result = {
        "metric": METRIC,
        "value": value,
        "unit": "nnz/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"powerlaw_mip single MIP m={inst.m} n={inst.n} nnz={inst.nnz} "
                               f"zipf_s={inst.meta['zipf_s']} max_row=50000 seed={args.seed}",
                   "sync": SYNC_NAMES[args.sync], "rounds": rounds, "global_batch": 1},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "mega_kernel" if args.sync == 2 else "round_kernel",
                     "kernel_ms": kernel_ms,
                     "bytes_per_launch": algorithmic_bytes(inst, rounds) if args.sync == 2 else None},
        "time_to_fixpoint_ms": ms_per_step,
    }
    if not args.no_cpu:
        import oracle_binding
        exp = oracle_binding.jacobi(inst, max_rounds=rounds)
        result["parity"] = {"bitwise_equal_to_oracle": bool(np.array_equal(out_lb, exp[0]) and
                                                            np.array_equal(out_ub, exp[1])),
                            "oracle_rounds": exp[2]}
        print("bench: timing cpu baselines", file=sys.stderr, flush=True)
        result["cpu_baseline"] = cpu_baseline(inst, args.cpu_runs)
        result["cpu_omp"] = omp_baseline(inst, min(16, os.cpu_count() or 8))
    prob.close()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
