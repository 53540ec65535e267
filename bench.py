#!/usr/bin/env python3
"""Propagation benchmark on one MI355X.

Workload: the single large sparse MIP (m = 1,000,000 rows, n = 500,000
columns, about 19.3M nonzeros, Zipf(1.95) row lengths in [2, 50,000]; see
gdp_amd.instances.powerlaw_mip), propagated to its fixpoint.  A step is one
propagation of the root bounds to the fixpoint: restore the root bounds on
the device (untimed), run all rounds.  The matrix and the bounds are resident
in HBM before the timed region starts.

Prints one JSON line:
  value         propagation throughput: nnz x rounds / second over the timed
                steps (kernel time of the propagations, HIP events)
  roofline      HBM roofline of the persistent kernel: algorithmic bytes per
                launch, rounds x (12 nnz + 20 m + 20 n), divided by the
                average launch time (HIP events on the library's stream around
                the launch), against 8 TB/s.  `traffic` is the measured HBM
                traffic per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE scaled by a
                copy kernel of known size), read from the committed profile
                summary when it was taken on this workload
  cpu_baseline  the oracle's sequential propagator (cpu_seq restated in C) on
                the same workload on this host, one thread

Other modes:
  --calibrate-copy MB   device-to-device copies of MB MiB (profiler
                        calibration of the byte counters), then exit
  --summarize DIR       read rocprofv3 outputs under DIR (kernel stats, FETCH_SIZE
                        and WRITE_SIZE passes, copy calibration passes) and write
                        the profile summary into profiles/
"""
import argparse
import csv
import glob
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
KERNEL_OF_SYNC = {0: "sweep_kernel", 1: "sweep_kernel", 2: "mega_kernel"}
PROFILE_JSON = os.path.join(REPO, "profiles", "mega_kernel_powerlaw_1m_rows.json")
PROFILE_MD = os.path.join(REPO, "profiles", "mega_kernel_powerlaw_1m_rows.md")


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
            "sample": f"cpu_seq (the oracle's C restatement: Gauss-Seidel rounds with row marking) on the full "
                      f"workload, median of {runs} runs; {rounds} rounds to the fixpoint in {t:.3f} s "
                      f"(value = nnz x its rounds / time)",
            "time_to_fixpoint_ms": t * 1e3, "rounds": rounds}


def omp_baseline(inst, threads):
    from gdp_amd import interface as I
    t0 = time.perf_counter()
    out = I.host_propagate(inst, "omp", threads=threads)
    t = time.perf_counter() - t0
    return {"threads": threads, "rounds": out.rounds, "time_to_fixpoint_ms": t * 1e3,
            "value": inst.nnz * out.rounds / t, "unit": "nnz/s"}


def load_traffic(inst, rounds, kernel):
    """Measured HBM bytes per launch from the committed profile summary, when
    it was taken on this workload and kernel; else None."""
    if not os.path.exists(PROFILE_JSON):
        return None
    with open(PROFILE_JSON) as f:
        p = json.load(f)
    w = p.get("workload", {})
    if (w.get("m"), w.get("n"), w.get("nnz"), w.get("rounds"), p.get("kernel")) != (inst.m, inst.n, inst.nnz, rounds,
                                                                                     kernel):
        return None
    return p.get("traffic_bytes_per_launch")


# ---------------------------------------------------------------------------
# profile summary (no GPU: reads rocprofv3 csv outputs)
# ---------------------------------------------------------------------------

def _csv_rows(pattern):
    files = sorted(glob.glob(pattern, recursive=True))
    if not files:
        raise SystemExit(f"no file matches {pattern}")
    rows = []
    for f in files:
        with open(f) as fh:
            rows.extend(csv.DictReader(fh))
    return rows


def _counter_per_dispatch(rows, name_part, counter):
    per = {}
    for r in rows:
        if name_part in r["Kernel_Name"] and r["Counter_Name"] == counter:
            d = int(r["Dispatch_Id"])
            per[d] = per.get(d, 0.0) + float(r["Counter_Value"])
    return per


def summarize(d, kernel="mega_kernel"):
    stats = _csv_rows(f"{d}/trace/**/*kernel_stats.csv")
    line = next(r for r in stats if kernel in r["Name"])
    avg_ns = float(line["AverageNs"])
    fetch = _counter_per_dispatch(_csv_rows(f"{d}/fetch/**/*counter_collection.csv"), kernel, "FETCH_SIZE")
    write = _counter_per_dispatch(_csv_rows(f"{d}/write/**/*counter_collection.csv"), kernel, "WRITE_SIZE")
    cal_f = _counter_per_dispatch(_csv_rows(f"{d}/cal_fetch/**/*counter_collection.csv"), "copyBuffer", "FETCH_SIZE")
    cal_w = _counter_per_dispatch(_csv_rows(f"{d}/cal_write/**/*counter_collection.csv"), "copyBuffer", "WRITE_SIZE")
    with open(f"{d}/meta.json") as fh:
        meta = json.load(fh)
    cal_bytes = meta["calibration_copy_bytes"]
    # the calibration copies are the largest copyBuffer dispatches of their run
    kf = statistics.median(sorted(cal_f.values())[-3:]) * 1024.0 / cal_bytes
    kw = statistics.median(sorted(cal_w.values())[-3:]) * 1024.0 / cal_bytes
    fetch_b = statistics.median(fetch.values()) * 1024.0
    write_b = statistics.median(write.values()) * 1024.0
    traffic = fetch_b / kf + write_b / kw
    alg = meta["algorithmic_bytes_per_launch"]
    # cross-check: read bytes from the L2-to-fabric read requests by size
    rdreq = None
    if glob.glob(f"{d}/rdreq/**/*counter_collection.csv", recursive=True):
        rr = _csv_rows(f"{d}/rdreq/**/*counter_collection.csv")
        sizes = {"TCC_EA0_RDREQ_32B_sum": 32, "TCC_EA0_RDREQ_64B_sum": 64, "TCC_EA0_RDREQ_128B_sum": 128}
        per = {}
        for c, b in sizes.items():
            for disp, v in _counter_per_dispatch(rr, kernel, c).items():
                per[disp] = per.get(disp, 0.0) + b * v
        rdreq = statistics.median(per.values())
    out = {
        "kernel": kernel,
        "workload": meta["workload"],
        "avg_launch_ns": avg_ns,
        "calls": int(line["Calls"]),
        "fetch_size_bytes_per_launch": fetch_b,
        "write_size_bytes_per_launch": write_b,
        "copy_calibration": {"bytes_copied": cal_bytes, "fetch_factor": kf, "write_factor": kw},
        "traffic_bytes_per_launch": traffic,
        "read_bytes_from_request_sizes": rdreq,
        "algorithmic_bytes_per_launch": alg,
        "achieved_gbs_algorithmic": alg / avg_ns,
        "achieved_gbs_traffic": traffic / avg_ns,
        "commands": meta["commands"],
    }
    os.makedirs(os.path.dirname(PROFILE_JSON), exist_ok=True)
    with open(PROFILE_JSON, "w") as fh:
        json.dump(out, fh, indent=1)
    w = meta["workload"]
    with open(PROFILE_MD, "w") as fh:
        fh.write(f"# {kernel} on the 1M-row power-law MIP\n\n")
        fh.write(f"Workload: m={w['m']} n={w['n']} nnz={w['nnz']}, {w['rounds']} rounds per launch "
                 f"(root bounds to the fixpoint).\n\n")
        fh.write("Commands (run from the repository root on one MI355X, in this order):\n\n")
        for c in meta["commands"]:
            fh.write(f"    {c}\n")
        fh.write("\n| quantity | value |\n|---|---|\n")
        fh.write(f"| {kernel} calls (kernel-trace run) | {int(line['Calls'])} |\n")
        fh.write(f"| average launch duration (rocprofv3 --stats) | {avg_ns / 1e3:.1f} us |\n")
        fh.write(f"| FETCH_SIZE per launch | {fetch_b / 1e6:.1f} MB |\n")
        fh.write(f"| WRITE_SIZE per launch | {write_b / 1e6:.1f} MB |\n")
        fh.write(f"| copy calibration ({cal_bytes / 2**20:.0f} MiB device copy) | FETCH x{kf:.3f}, WRITE x{kw:.3f} |\n")
        fh.write(f"| HBM traffic per launch (calibrated) | {traffic / 1e6:.1f} MB |\n")
        if rdreq is not None:
            fh.write(f"| read bytes from 32/64/128-byte read request counts (cross-check of the FETCH factor) | "
                     f"{rdreq / 1e6:.1f} MB |\n")
        fh.write(f"| algorithmic bytes per launch | {alg / 1e6:.1f} MB |\n")
        fh.write(f"| achieved, algorithmic bytes / duration | {alg / avg_ns:.1f} GB/s |\n")
        fh.write(f"| achieved, measured traffic / duration | {traffic / avg_ns:.1f} GB/s |\n")
    print(json.dumps(out))


# ---------------------------------------------------------------------------

def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--sync", type=int, default=2, choices=[0, 1, 2])
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-runs", type=int, default=7)
    ap.add_argument("--no-cpu", action="store_true", help="skip the host baselines and the parity check")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MB")
    ap.add_argument("--summarize", default=None, metavar="DIR")
    ap.add_argument("--meta-out", default=None, help="write workload facts for --summarize")
    args = ap.parse_args()
    if args.summarize:
        summarize(args.summarize)
        return
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
    if args.calibrate_copy:
        nb = args.calibrate_copy << 20
        src = I.DeviceArray(np.zeros(nb // 8))
        dst = I.DeviceArray(np.zeros(nb // 8))
        for _ in range(3):
            dst.copy_from(src)
        print(json.dumps({"calibration_copy_bytes": nb}), flush=True)
        return

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
        lb.copy_from(root_lb)   # synchronous device-to-device copies, outside the kernel timing
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
    kernel_ms = statistics.mean(i.kernel_ms for i in infos)
    ms_per_step = kernel_ms
    value = inst.nnz * rounds / (kernel_ms * 1e-3)
    alg = algorithmic_bytes(inst, rounds)
    achieved = alg / (kernel_ms * 1e-3) / 1e9
    wall_ms = (t1 - t0) / args.steps * 1e3
    kernel = KERNEL_OF_SYNC[args.sync]
    print(f"bench: {rounds} rounds, kernel {kernel_ms:.3f} ms/step, wall {wall_ms:.3f} ms/step incl. bound restore, "
          f"{achieved:.1f} GB/s algorithmic", file=sys.stderr, flush=True)
    if args.meta_out:
        with open(args.meta_out, "w") as fh:
            json.dump({"workload": {"m": inst.m, "n": inst.n, "nnz": inst.nnz, "rounds": rounds},
                       "algorithmic_bytes_per_launch": alg}, fh)

    out_lb, out_ub = lb.numpy(), ub.numpy()
    traffic = load_traffic(inst, rounds, kernel) if args.sync == 2 else None
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
                               f"zipf_s={inst.meta['zipf_s']} row_len=[2,50000] seed={args.seed}, "
                               f"root bounds to the fixpoint",
                   "sync": SYNC_NAMES[args.sync], "rounds": rounds, "global_batch": 1},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": kernel, "launch_ms": kernel_ms, "algorithmic_bytes_per_launch": alg},
        "wall_ms_per_step": wall_ms,
    }
    if not args.no_cpu:
        import oracle_binding
        print("bench: parity check against the oracle", file=sys.stderr, flush=True)
        exp = oracle_binding.jacobi(inst, max_rounds=100)
        result["parity"] = {"bitwise_equal_to_oracle": bool(np.array_equal(out_lb, exp[0]) and
                                                            np.array_equal(out_ub, exp[1])),
                            "rounds_equal": exp[2] == rounds}
        print("bench: timing cpu baselines", file=sys.stderr, flush=True)
        result["cpu_baseline"] = cpu_baseline(inst, args.cpu_runs)
        result["cpu_omp"] = omp_baseline(inst, min(16, os.cpu_count() or 8))
    prob.close()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
