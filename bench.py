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
                launch, summed over the rounds of 12 nnz_r + 20 m_r + 20 n_r
                for the rows a round processes (row marking skips rows none of
                whose columns changed; n_r: distinct columns of those rows,
                from an untimed diagnostic run of the same propagation),
                divided by the average launch time (HIP events on the
                library's stream around the launch), against 8 TB/s; the
                dense-sweep equivalent, rounds x (12 nnz + 20 m + 20 n), is
                reported beside it.  `traffic` is the measured HBM
                traffic per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE scaled by a
                copy kernel of known size), read from the committed profile
                summary when it was taken on this workload
  cpu_baseline  the oracle's sequential propagator (cpu_seq restated in C) on
                the same workload on this host, one thread
  batched       256 independent MIPs of uneven size (m log-uniform in
                [1e3, 2e5], n = m/2, mean row length log-uniform in [3, 30]) in
                shared launches vs back-to-back plain launches, instances/s;
                every instance checked against its own cpu_seq
  serving       node requests against a device-resident 2M-nonzero matrix:
                1024 nodes in one shared launch, and a continuous-batching
                server at Poisson arrival rates 1k..50k/s with p50/p99
                latency, vs serial plain launches; requests as full bound
                vectors (copied through the server's staging), as full
                vectors in caller-owned pinned memory (zero-copy; results in
                separate pinned arrays, or in place with only the changed
                columns coming back), and as changes to the propagated root

Other modes:
  --calibrate-copy MB   device-to-device copies of MB MiB (profiler
                        calibration of the byte counters), then exit
  --summarize DIR       read rocprofv3 outputs under DIR (kernel stats, FETCH_SIZE
                        and WRITE_SIZE passes, copy calibration passes) and write
                        the profile summary into profiles/
"""
import argparse
import csv
import ctypes
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
    """Bytes of `rounds` dense sweeps: per nonzero its column index and
    coefficient, per row its extent and sides, per column its bounds and type."""
    return rounds * (12 * inst.nnz + 20 * inst.m + 20 * inst.n)


def processed_bytes(rows, nnz, cols):
    """Same per-unit bytes over the rows each round actually processes (and
    the distinct columns those rows hold)."""
    return int(12 * sum(nnz) + 20 * sum(rows) + 20 * sum(cols))


def host_facts():
    """The host this run's CPU baselines ran on."""
    model, cores = None, set()
    phys = core = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name" and model is None:
                    model = v
                elif k == "physical id":
                    phys = v
                elif k == "core id":
                    core = v
                elif not k and phys is not None:
                    cores.add((phys, core))
                    phys = core = None
        if phys is not None:
            cores.add((phys, core))
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count()
    return {"nproc": os.cpu_count(), "usable_cpus": usable, "model_name": model,
            "physical_cores": len(cores) or None, "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS")}


def cpu_baseline(inst, runs):
    """cpu_seq (the oracle's C restatement), one thread, timed like the
    reference: the round loop alone, without the column index build and the
    allocations (sequential_propagator.h:248-281)."""
    import oracle_binding
    loops, walls, rounds, res = [], [], 0, None
    for _ in range(runs):
        t0 = time.perf_counter()
        res, loop_s = oracle_binding.seq_timed(inst)
        walls.append(time.perf_counter() - t0)
        loops.append(loop_s)
        rounds = res[2]
    t = statistics.median(loops)
    return {"value": inst.nnz * rounds / t, "unit": "nnz/s", "cores": 1, "kind": "port",
            "sample": f"cpu_seq (the oracle's C restatement: Gauss-Seidel rounds with row marking) on the full "
                      f"workload, median of {runs} runs of the round loop alone (no CSC build or allocation, as the "
                      f"reference times it); {rounds} rounds to the fixpoint in {t:.3f} s "
                      f"(value = nnz x its rounds / time)",
            "time_to_fixpoint_ms": t * 1e3, "wall_with_setup_ms": statistics.median(walls) * 1e3,
            "rounds": rounds}, res


def deviation_vs_seq(inst, lbs, ubs, seq):
    """Integer mismatches and continuous deviation (relative or absolute) of
    the GPU fixpoint against cpu_seq's."""
    import numpy as np
    is_int = inst.vartype != 3
    bad_int = int(np.sum(lbs[is_int] != seq[0][is_int]) + np.sum(ubs[is_int] != seq[1][is_int]))
    worst, beyond = 0.0, 0
    for g, sv in ((lbs, seq[0]), (ubs, seq[1])):
        d = np.abs(g - sv)[~is_int] / np.maximum(1.0, np.abs(sv[~is_int]))
        d[g[~is_int] == sv[~is_int]] = 0.0
        worst = max(worst, float(d.max()) if d.size else 0.0)
        beyond += int(np.sum(d > 1e-9))
    return bad_int, worst, beyond


def batched_parity(out):
    """Every batched instance against the round-synchronous oracle (bit for
    bit: bounds, rounds, flags) and, where neither run hit the round cap,
    against its own cpu_seq fixpoint: integer bounds exact, continuous within
    1e-9 relative or absolute, and each fixpoint one of the other
    propagator (a round of either started from the other's accepts no
    candidate)."""
    import numpy as np
    import oracle_binding
    insts, o, out_l, out_u, infos = out["insts"], out["offsets"], out["lb"], out["ub"], out["infos"]
    exact, bad_int, worst, feas, compared, capped, outside, mutual = 0, 0, 0.0, 0, 0, [], [], 0
    for i, x in enumerate(insts):
        gl, gu = out_l[o[i]:o[i + 1]], out_u[o[i]:o[i + 1]]
        rounds, infeasible, is_capped = infos[i]
        jac = oracle_binding.jacobi(x)
        exact += bool(np.array_equal(gl, jac[0]) and np.array_equal(gu, jac[1]) and jac[2] == rounds and
                      jac[3] == infeasible and jac[4] == is_capped)
        seq = oracle_binding.seq(x)
        feas += infeasible == bool(seq[3])
        if is_capped or seq[4] or seq[3]:
            if is_capped or seq[4]:
                capped.append(i)
            continue
        compared += 1
        b, w, beyond = deviation_vs_seq(x, gl, gu, seq)
        bad_int += b
        worst = max(worst, w)
        if beyond:
            outside.append({"instance": i, "bounds": beyond, "max_dev": w})
        again = oracle_binding.seq(x, lb=gl, ub=gu)
        back = oracle_binding.jacobi(x, lb=seq[0], ub=seq[1])
        mutual += bool(again[2] == 1 and np.array_equal(again[0], gl) and np.array_equal(again[1], gu) and
                       back[2] == 1 and np.array_equal(back[0], seq[0]) and np.array_equal(back[1], seq[1]))
    return {"bitwise_equal_to_oracle": exact, "of": len(insts),
            "vs_cpu_seq": {"compared": compared, "capped_at_100_rounds": capped,
                           "integer_bound_mismatches": bad_int,
                           "max_continuous_rel_or_abs_dev": worst,
                           "instances_with_continuous_bounds_beyond_1e-9": outside,
                           "mutual_epsilon_fixpoints": mutual,
                           "infeasibility_agrees": feas == len(insts)}}


def cpu_leg(inst, out_lb, out_ub, rounds, runs, batched_out):
    """The host side: cpu_seq timed as the baseline, and the oracle as the
    checker of the GPU results of the other legs."""
    import numpy as np
    import oracle_binding
    print("bench: parity check against the oracle", file=sys.stderr, flush=True)
    exp = oracle_binding.jacobi(inst, max_rounds=100)
    parity = {"bitwise_equal_to_oracle": bool(np.array_equal(out_lb, exp[0]) and np.array_equal(out_ub, exp[1])),
              "rounds_equal": exp[2] == rounds}
    print("bench: timing cpu baselines", file=sys.stderr, flush=True)
    base, seq = cpu_baseline(inst, runs)
    bad_int, worst, beyond = deviation_vs_seq(inst, out_lb, out_ub, seq)
    parity["vs_cpu_seq"] = {"integer_bound_mismatches": bad_int, "max_continuous_rel_or_abs_dev": worst,
                            "continuous_bounds_beyond_1e-9": beyond, "infeasibility_agrees": bool(seq[3]) is False,
                            "rounds_cpu_seq": seq[2], "rounds_gpu": rounds}
    print(f"bench: vs cpu_seq: {bad_int} integer mismatches, max continuous deviation {worst:.3e}",
          file=sys.stderr, flush=True)
    batched = batched_parity(batched_out) if batched_out is not None else None
    return parity, base, batched


def omp_baseline(inst, threads):
    """The library's cpu_omp; time of the round loop alone (wall time with
    the column index build beside it)."""
    from gdp_amd import interface as I
    t0 = time.perf_counter()
    out = I.host_propagate(inst, "omp", threads=threads)
    t = time.perf_counter() - t0
    return {"threads": threads, "rounds": out.rounds, "time_to_fixpoint_ms": out.kernel_ms,
            "wall_with_setup_ms": t * 1e3, "value": inst.nnz * out.rounds / (out.kernel_ms * 1e-3), "unit": "nnz/s"}


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
# batched and served propagation
# ---------------------------------------------------------------------------

def batched_leg(sync):
    """256 uneven MIPs: one shared run (with the default tuning, then with
    the autotuner's pick) vs back-to-back plain launches.  Returns the result
    record and the outputs for the parity check."""
    import numpy as np
    from gdp_amd import interface as I
    from gdp_amd.instances import batch_family, small_mip
    insts = [small_mip(**sp) for sp in batch_family(count=256)]
    nnz = sum(len(x.col) for x in insts)
    print(f"bench: batched workload, {len(insts)} instances, {nnz} nonzeros", file=sys.stderr, flush=True)
    b = I.Batch(insts)
    lb, ub = b.concat_bounds()
    root_l, root_u = I.DeviceArray(lb), I.DeviceArray(ub)
    dl, du = I.DeviceArray(lb), I.DeviceArray(ub)
    times = []
    for rep in range(4):
        dl.copy_from(root_l)
        du.copy_from(root_u)
        infos = b.propagate_device(dl.ptr, du.ptr, sync)
        if rep:
            times.append(infos[0].kernel_ms)
    default_ms = statistics.median(times)
    out_l, out_u = dl.numpy(), du.numpy()
    b.close()
    # autotuned layout and persistent grid: a coordinate search on a sample of
    # the batch (every 8th instance), applied to the whole batch
    print("bench: autotuning the batch layout on a sample of 32 instances", file=sys.stderr, flush=True)
    t_tune = time.perf_counter()
    tuning, _ = I.batch_autotune(insts[::8], sync)
    tune_s = time.perf_counter() - t_tune
    b = I.Batch(insts, tuning)
    times = []
    for rep in range(4):
        dl.copy_from(root_l)
        du.copy_from(root_u)
        tuned_infos = b.propagate_device(dl.ptr, du.ptr, sync)
        if rep:
            times.append(tuned_infos[0].kernel_ms)
    tuned_ms = statistics.median(times)
    assert np.array_equal(dl.numpy(), out_l) and np.array_equal(du.numpy(), out_u), "tuned batch differs"
    batched_ms = tuned_ms
    # back-to-back plain launches: one device-resident problem per instance
    probs = [I.Problem(x) for x in insts]
    arrs = [(I.DeviceArray(x.lb), I.DeviceArray(x.ub)) for x in insts]
    plain_ms = 0.0
    for p, (a, c), x in zip(probs, arrs, insts):
        info = p.propagate_device(a.ptr, c.ptr, sync)
        plain_ms += info.kernel_ms
    res = {"workload": f"256 small_mip instances, m log-uniform [1e3, 2e5], n = m/2, mean row length log-uniform "
                       f"[3, 30], {nnz} nonzeros in total, root bounds to the fixpoint",
           "sync": SYNC_NAMES[sync], "instances": len(insts), "nnz_total": nnz,
           "batched_ms": batched_ms, "nodes_per_s": len(insts) / (batched_ms * 1e-3),
           "default_tuning_ms": default_ms, "autotuned": tuning, "autotune_s": tune_s,
           "plain_ms": plain_ms, "plain_nodes_per_s": len(insts) / (plain_ms * 1e-3),
           "speedup": plain_ms / batched_ms, "max_rounds": max(i.rounds for i in infos)}
    for p in probs:
        p.close()
    outputs = {"insts": insts, "offsets": list(b.n_offsets), "lb": out_l, "ub": out_u,
               "infos": [(i.rounds, bool(i.infeasible), bool(i.capped)) for i in infos]}
    b.close()
    return res, outputs


def serving_leg(sync):
    """Node requests against one device-resident matrix."""
    import numpy as np
    from gdp_amd import interface as I
    from gdp_amd.instances import powerlaw_mip, tighten_node
    inst = powerlaw_mip(seed=4, m=200_000, n=100_000, s=2.08, len_hi=5000)
    print(f"bench: serving workload, nnz={inst.nnz}", file=sys.stderr, flush=True)
    p = I.Problem(inst)
    root = p.propagate(inst.lb, inst.ub, sync)
    rng = np.random.default_rng(7)
    pool = [tighten_node(inst, rng, 1, 50, root.lbs, root.ubs) for _ in range(400)]
    t = time.perf_counter()
    rounds = [p.propagate(lb, ub, sync).rounds for lb, ub, _ in pool[:200]]
    plain = 200 / (time.perf_counter() - t)
    B = 1024
    p.reserve(B)
    dls = [I.DeviceArray(pool[k % len(pool)][0]) for k in range(B)]
    dus = [I.DeviceArray(pool[k % len(pool)][1]) for k in range(B)]
    hints = [pool[k % len(pool)][2] for k in range(B)]
    for rep in range(2):
        for k in range(B):
            dls[k].upload(pool[k % len(pool)][0])
            dus[k].upload(pool[k % len(pool)][1])
        infos = p.propagate_nodes_device([d.ptr for d in dls], [d.ptr for d in dus], hints, sync)
    shared = B / (infos[0].kernel_ms * 1e-3)
    for d in dls + dus:
        d.free()
    srv = I.Server(p, hbm_budget=16 << 30, max_slots=1024)
    srv.set_base(root.lbs, root.ubs)
    # The load generator (gpu-domain-propagator_amd/bench/serve_load.cpp) is
    # native: 4 submitting threads issue the requests at their Poisson
    # arrival times and one thread takes the results, through the C API,
    # while the server steps on its own thread.
    load = ctypes.CDLL(os.path.join(REPO, "gpu-domain-propagator_amd", "lib", "libgdp_serveload.so"))
    load.gdpServeLoad.restype = ctypes.c_int
    dp = ctypes.POINTER(ctypes.c_double)
    pool_l = (dp * len(pool))(*[x[0].ctypes.data_as(dp) for x in pool])
    pool_u = (dp * len(pool))(*[x[1].ctypes.data_as(dp) for x in pool])
    cols = [np.ascontiguousarray(x[2], np.int32) for x in pool]
    ip = ctypes.POINTER(ctypes.c_int)
    pool_c = (ip * len(pool))(*[c.ctypes.data_as(ip) for c in cols])
    pool_n = (ctypes.c_int * len(pool))(*[len(c) for c in cols])
    srv.start()
    rates = []
    modes = {"full": 0, "pinned": 1, "inplace": 2, "delta": 3}
    for mode in ("full", "pinned", "inplace", "delta"):
        for rate in (1000, 2000, 5000, 10000, 20000, 50000):
            N = int(max(1000, min(10000, rate // 4)))
            arr = np.cumsum(rng.exponential(1.0 / rate, size=N))
            arr -= arr[0]
            done = np.zeros(N)
            ok = load.gdpServeLoad(srv.handle, modes[mode], N, arr.ctypes.data_as(dp), len(pool), inst.n, pool_l,
                                   pool_u, pool_c, pool_n, 4, 256, done.ctypes.data_as(dp), None)
            if not ok:
                raise RuntimeError(f"serving load {mode} {rate}/s failed: {I.last_error()}")
            lat = (done - arr) * 1e3
            rates.append({"requests": mode, "offered_per_s": rate, "n": N,
                          "achieved_per_s": N / done.max(),
                          "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99))})
            print(f"bench: serving {mode} {rate}/s -> {rates[-1]['achieved_per_s']:.0f}/s "
                  f"p50 {rates[-1]['p50_ms']:.2f} ms p99 {rates[-1]['p99_ms']:.2f} ms", file=sys.stderr, flush=True)
    srv.stop()
    srv.close()
    p.close()
    plateau = {}
    for mode in modes:
        rs = [r for r in rates if r["requests"] == mode]
        top = max(r["achieved_per_s"] for r in rs)
        at_max = next(r for r in rs if r["offered_per_s"] == 50000)["achieved_per_s"]
        plateau[mode] = {"best_per_s": top, "at_50k_offered_per_s": at_max, "ratio": at_max / top}
    best = max(r["achieved_per_s"] for r in rates)
    best_full = max(r["achieved_per_s"] for r in rates if r["requests"] in ("full", "pinned", "inplace"))
    return {"workload": f"powerlaw_mip m={inst.m} n={inst.n} nnz={inst.nnz} (zipf 2.08, rows <= 5000); nodes = "
                        f"propagated root with 1..50 integer columns restricted to half their domain",
            "sync": SYNC_NAMES[sync], "slots": srv.slots, "mean_node_rounds": float(np.mean(rounds)),
            "plain_nodes_per_s": plain, "shared_launch_1024_nodes_per_s": shared,
            "shared_speedup": shared / plain, "served_best_per_s": best, "served_speedup": best / plain,
            "served_full_vectors_best_per_s": best_full, "served_full_vectors_speedup": best_full / plain,
            "plateau_under_overload": plateau, "rates": rates}


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


PROFILE_TIMED = "--steps 10 --warmup 2 --no-cpu --no-batch --no-serve --no-extra --no-tune"
PROFILE_PASS = "--steps 2 --warmup 1 --no-cpu --no-batch --no-serve --no-extra --no-tune"
PROFILE_RDREQ = "TCC_EA0_RDREQ_32B_sum TCC_EA0_RDREQ_64B_sum TCC_EA0_RDREQ_128B_sum TCC_EA0_RDREQ_sum"
PROFILE_CAL_MB = 256


def profile_commands(d):
    """The runs whose outputs under `d` summarize() reads, each rocprofv3 run
    in a process of its own."""
    return [
        "cd /tmp && export TMPDIR=/tmp && cd <repository root>  (each rocprofv3 run on its own)",
        f"rocprofv3 --kernel-trace --stats --output-format csv -d {d}/trace -o bench -- python3 bench.py "
        f"{PROFILE_TIMED} --meta-out {d}/bench_meta.json",
        f"rocprofv3 --pmc FETCH_SIZE --output-format csv -d {d}/fetch -o bench -- python3 bench.py {PROFILE_PASS}",
        f"rocprofv3 --pmc WRITE_SIZE --output-format csv -d {d}/write -o bench -- python3 bench.py {PROFILE_PASS}",
        f"rocprofv3 --pmc FETCH_SIZE --output-format csv -d {d}/cal_fetch -o cal -- python3 bench.py "
        f"--calibrate-copy {PROFILE_CAL_MB}",
        f"rocprofv3 --pmc WRITE_SIZE --output-format csv -d {d}/cal_write -o cal -- python3 bench.py "
        f"--calibrate-copy {PROFILE_CAL_MB}",
        f"rocprofv3 --pmc {PROFILE_RDREQ} --output-format csv -d {d}/rdreq -o bench -- python3 bench.py {PROFILE_PASS}",
        f"rocprofv3 --pmc {PROFILE_RDREQ} --output-format csv -d {d}/cal_rdreq -o cal -- python3 bench.py "
        f"--calibrate-copy {PROFILE_CAL_MB}",
        f"python3 bench.py --summarize {d}",
    ]


def summarize(d, kernel="mega_kernel"):
    stats = _csv_rows(f"{d}/trace/**/*kernel_stats.csv")
    line = next(r for r in stats if kernel in r["Name"])
    # workload facts written by the kernel-trace run (--meta-out)
    with open(f"{d}/bench_meta.json") as fh:
        meta = json.load(fh)
    meta["calibration_copy_bytes"] = PROFILE_CAL_MB << 20
    meta["commands"] = profile_commands(d)
    # the timed launches: the kernel trace in dispatch order, less the warmup
    trace = [r for r in _csv_rows(f"{d}/trace/**/*kernel_trace.csv") if kernel in r["Kernel_Name"]]
    trace.sort(key=lambda r: int(r["Start_Timestamp"]))
    durs = [int(r["End_Timestamp"]) - int(r["Start_Timestamp"]) for r in trace]
    timed = durs[meta.get("warmup", 0):][:meta.get("steps", len(durs))]
    avg_ns = statistics.mean(timed)
    fetch = _counter_per_dispatch(_csv_rows(f"{d}/fetch/**/*counter_collection.csv"), kernel, "FETCH_SIZE")
    write = _counter_per_dispatch(_csv_rows(f"{d}/write/**/*counter_collection.csv"), kernel, "WRITE_SIZE")
    cal_f = _counter_per_dispatch(_csv_rows(f"{d}/cal_fetch/**/*counter_collection.csv"), "copyBuffer", "FETCH_SIZE")
    cal_w = _counter_per_dispatch(_csv_rows(f"{d}/cal_write/**/*counter_collection.csv"), "copyBuffer", "WRITE_SIZE")
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
        "timed_launches": len(timed),
        "stats_average_ns_all_launches": float(line["AverageNs"]),
        "bench_launch_ms_same_run": meta.get("bench_launch_ms"),
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
        fh.write(f"| average duration of the {len(timed)} timed launches (kernel trace, warmup excluded) | "
                 f"{avg_ns / 1e3:.1f} us |\n")
        fh.write(f"| rocprofv3 --stats average over all {int(line['Calls'])} launches (incl. warmup) | "
                 f"{float(line['AverageNs']) / 1e3:.1f} us |\n")
        if meta.get("bench_launch_ms"):
            fh.write(f"| bench.py's own HIP-event average of the same launches | {meta['bench_launch_ms'] * 1e3:.1f} us |\n")
        fh.write(f"| FETCH_SIZE per launch | {fetch_b / 1e6:.1f} MB |\n")
        fh.write(f"| WRITE_SIZE per launch | {write_b / 1e6:.1f} MB |\n")
        fh.write(f"| copy calibration ({cal_bytes / 2**20:.0f} MiB device copy) | FETCH x{kf:.3f}, WRITE x{kw:.3f} |\n")
        fh.write(f"| HBM traffic per launch (calibrated) | {traffic / 1e6:.1f} MB |\n")
        if rdreq is not None:
            fh.write(f"| read bytes from 32/64/128-byte read request counts (cross-check of the FETCH factor) | "
                     f"{rdreq / 1e6:.1f} MB |\n")
        fh.write(f"| algorithmic bytes per launch (rows each round processes) | {alg / 1e6:.1f} MB |\n")
        fh.write(f"| achieved, algorithmic bytes / duration | {alg / avg_ns:.1f} GB/s |\n")
        fh.write(f"| achieved, measured traffic / duration | {traffic / avg_ns:.1f} GB/s |\n")
    print(json.dumps(out))


def summarize_steps(d, out_md):
    """Per-round durations of the step kernels (sweep, apply, publish) of the
    CPU_LOOP propagations in a rocprofv3 kernel trace under d: one row per
    round of the last propagation of each round count (the timed ones come
    after the warmup), written as a markdown table to out_md."""
    trace = _csv_rows(f"{d}/**/*kernel_trace.csv")
    trace.sort(key=lambda r: int(r["Start_Timestamp"]))
    props, cur = [], None
    for r in trace:
        name = r["Kernel_Name"].split("(")[0].split("::")[-1]
        dur = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
        if name == "load_kernel":
            cur = {"steps": [], "start": int(r["Start_Timestamp"])}
        elif cur is None:
            continue
        elif name == "sweep_kernel":
            cur["steps"].append([dur, 0.0, 0.0])
        elif name == "apply_split_kernel" and cur["steps"]:
            cur["steps"][-1][1] = dur
        elif name == "publish_kernel" and cur["steps"]:
            cur["steps"][-1][2] = dur
        elif name == "unpack_kernel":
            cur["wall"] = (int(r["End_Timestamp"]) - cur["start"]) / 1e3
            props.append(cur)
            cur = None
    last = {}  # the fastest propagation of each round count (a timed one)
    for p in props:
        k = len(p["steps"])
        if k not in last or p["wall"] < last[k]["wall"]:
            last[k] = p
    with open(out_md, "w") as fh:
        fh.write("# Step kernels per round (CPU_LOOP launches)\n\n")
        fh.write("Commands (from the repository root, `cd /tmp && export TMPDIR=/tmp` first):\n\n")
        fh.write(f"    rocprofv3 --kernel-trace --stats --output-format csv -d {d} -o steps -- python3 bench.py "
                 "--sync 0 --steps 3 --warmup 1 --no-cpu --no-batch --no-serve --no-extra --no-tune\n")
        fh.write(f"    python3 bench.py --summarize-steps {d} {out_md}\n\n")
        fh.write("Durations in microseconds from the kernel trace, one propagation per round count "
                 "(the fastest one from its load to its unpack: a timed step).\n\n")
        for k in sorted(last):
            p = last[k]
            tot = [sum(x[i] for x in p["steps"]) for i in range(3)]
            fh.write(f"## {k} rounds: sweep {tot[0]:.0f} us, apply {tot[1]:.0f} us, publish {tot[2]:.0f} us, "
                     f"load to unpack {p['wall']:.0f} us\n\n| round | sweep | apply | publish |\n|---|---|---|---|\n")
            for i, x in enumerate(p["steps"]):
                fh.write(f"| {i + 1} | {x[0]:.1f} | {x[1]:.1f} | {x[2]:.1f} |\n")
            fh.write("\n")
    print(f"wrote {out_md}")


def summarize_batch(d, out_md):
    """The batched leg from a rocprofv3 kernel trace under d of a bench run
    whose stdout (the JSON line) is in d/bench.out: the last 260 mega_kernel
    dispatches are the 4 launches of the tuned batch and the 256 plain
    launches, in that order."""
    trace = [r for r in _csv_rows(f"{d}/**/*kernel_trace.csv") if "mega_kernel" in r["Kernel_Name"]]
    trace.sort(key=lambda r: int(r["Start_Timestamp"]))
    dur = [(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e6 for r in trace]
    grid = [int(r.get("Grid_Size_X") or r.get("Grid_Size") or 0) // max(1, int(r.get("Workgroup_Size_X") or
                                                                             r.get("Workgroup_Size") or 1))
            for r in trace]
    wg = [int(r.get("Workgroup_Size_X") or r.get("Workgroup_Size") or 0) for r in trace]
    if len(dur) < 260:
        raise SystemExit(f"{len(dur)} mega_kernel dispatches, expected at least 260")
    batch, plain = dur[-260:-256], dur[-256:]
    line = None
    with open(f"{d}/bench.out") as fh:
        for ln in fh:
            if ln.startswith("{") and '"metric"' in ln:
                line = json.loads(ln)
    bt = line["batched"]
    tuned = statistics.median(batch[1:])
    with open(out_md, "w") as fh:
        fh.write("# mega_kernel on the batch of 256 MIPs, against 256 plain launches\n\n")
        fh.write(f"Workload: {bt['workload']}. All instances share one persistent launch of the library's batch; "
                 "the baseline runs one launch per instance, back to back, each on its own device-resident "
                 "problem.\n\n")
        fh.write("Commands (from the repository root on one MI355X, `cd /tmp && export TMPDIR=/tmp` first):\n\n")
        fh.write(f"    rocprofv3 --kernel-trace --stats --output-format csv -d {d} -o batch -- python3 bench.py "
                 f"--no-serve --no-cpu --no-extra --no-tune > {d}/bench.out\n")
        fh.write(f"    python3 bench.py --summarize-batch {d} {out_md}\n\n")
        fh.write(f"The kernel trace holds {len(dur)} `mega_kernel` dispatches in the bench's order (the 1M-row "
                 "problem, the drop-in call, the batch with the default layout, the autotuner's trials on a "
                 "sample of 32 instances, the batch with the tuned layout, the 256 plain launches); the figures "
                 "are the last 260 (end minus start timestamp).\n\n")
        fh.write("| quantity | value |\n|---|---|\n")
        fh.write(f"| tuned batch, 4 launches | {', '.join(f'{x:.3f}' for x in batch)} ms |\n")
        fh.write(f"| tuned batch, median of the 3 timed launches (the first warms up) | {tuned:.2f} ms |\n")
        fh.write(f"| bench.py's own figure for the same launches (HIP events, median) | {bt['batched_ms']:.2f} ms |\n")
        fh.write(f"| grid of the batch launch | {grid[-257]} workgroups x {wg[-257]} threads |\n")
        fh.write(f"| 256 plain launches, sum of kernel durations | {sum(plain):.1f} ms ({min(plain):.2f} ms to "
                 f"{max(plain):.2f} ms each) |\n")
        fh.write(f"| bench.py's own sum for the plain launches (HIP events around each) | {bt['plain_ms']:.1f} ms |\n")
        fh.write(f"| grids of the plain launches | from {min(grid[-256:])} to {max(grid[-256:])} workgroups, sized "
                 "to each instance |\n")
        fh.write(f"| speedup, kernel trace | {sum(plain):.1f} / {tuned:.2f} = {sum(plain) / tuned:.1f}x |\n")
        fh.write(f"| speedup, bench.py | {bt['plain_ms']:.1f} / {bt['batched_ms']:.2f} = {bt['speedup']:.1f}x |\n\n")
        fh.write("`bench.py` asserts that the tuned layout's results equal the default layout's; its full run "
                 "(without `--no-cpu`) compares all 256 results with the oracle bit for bit.\n")
    print(f"wrote {out_md}")


# ---------------------------------------------------------------------------

def measured_copy_gbs(nbytes=1 << 30, reps=5):
    """HBM bandwidth this device reaches in a plain device-to-device copy of
    nbytes: (read + write bytes) / time, best of reps (STREAM-style copy)."""
    import numpy as np
    from gdp_amd import interface as I
    src = I.DeviceArray(np.zeros(nbytes // 8))
    dst = I.DeviceArray(np.zeros(nbytes // 8))
    dst.copy_from(src)
    best = float("inf")
    for _ in range(reps):
        t = time.perf_counter()
        dst.copy_from(src)  # synchronous
        best = min(best, time.perf_counter() - t)
    src.free()
    dst.free()
    return 2.0 * nbytes / best / 1e9


def workload_leg(inst, sync, steps, label, fixed_cost_rounds=0):
    """One more workload propagated to its fixpoint: kernel time per
    propagation (mean over `steps` after a warmup, inputs resident), rounds,
    nnz x rounds / s and the bytes of the rows each round processes over the
    kernel time.  With fixed_cost_rounds = k, also the time of each of the
    last k rounds, from runs capped at R - k, ..., R rounds (R: the rounds to
    the fixpoint) and differenced: those rounds process few rows, so their
    time is mostly the fixed cost of a round (barriers, publish, item scan)."""
    from gdp_amd import interface as I
    print(f"bench: {label}: m={inst.m} n={inst.n} nnz={inst.nnz}", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    prob = I.Problem(inst)
    create_ms = (time.perf_counter() - t0) * 1e3
    root_l, root_u = I.DeviceArray(inst.lb), I.DeviceArray(inst.ub)
    dl, du = I.DeviceArray(inst.lb), I.DeviceArray(inst.ub)

    def run(max_rounds=100):
        dl.copy_from(root_l)
        du.copy_from(root_u)
        return prob.propagate_device(dl.ptr, du.ptr, sync, max_rounds)

    run()
    infos = [run() for _ in range(steps)]
    rounds = infos[0].rounds
    ms = statistics.mean(i.kernel_ms for i in infos)
    wp, w_rows, w_nnz, w_cols, _ = prob.work_profile(inst.lb, inst.ub)
    alg = processed_bytes(w_rows, w_nnz, w_cols)
    out = {"workload": label, "m": inst.m, "n": inst.n, "nnz": inst.nnz, "sync": SYNC_NAMES[sync],
           "rounds": rounds, "infeasible": bool(infos[0].infeasible), "kernel_ms": ms, "ms_per_round": ms / rounds,
           "nnz_rounds_per_s": inst.nnz * rounds / (ms * 1e-3), "problem_create_ms": create_ms,
           "processed_bytes_per_launch": alg, "achieved_gbs": alg / (ms * 1e-3) / 1e9,
           "frac": alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
           "rows_processed_per_round": [int(x) for x in w_rows]}
    if fixed_cost_rounds and rounds > fixed_cost_rounds + 1:
        caps = list(range(rounds - fixed_cost_rounds - 1, rounds + 1))
        t = {c: statistics.median(run(c).kernel_ms for _ in range(3)) for c in caps}
        per = [(t[c] - t[c - 1]) * 1e3 for c in caps[1:]]
        out["last_rounds_us"] = per
        out["last_rounds_rows"] = [int(w_rows[c - 1]) for c in caps[1:]]
        out["fixed_round_cost_us"] = statistics.median(per)
    for a in (root_l, root_u, dl, du):
        a.free()
    prob.close()
    print(f"bench: {label}: {rounds} rounds, {ms:.3f} ms, {out['ms_per_round'] * 1e3:.1f} us/round"
          + (f", last rounds {out['fixed_round_cost_us']:.1f} us each" if "fixed_round_cost_us" in out else ""),
          file=sys.stderr, flush=True)
    return out


def end_to_end_leg(inst, out_lb, out_ub, sync, runs=3):
    """The drop-in call as a caller of the reference's C interface makes it
    (propagateConstraintsGPUAtomicDouble): host arrays in, host arrays out,
    the device layout built and freed inside the call.  Best of `runs`; the
    result is checked bit for bit against the timed propagation."""
    import numpy as np
    from gdp_amd import interface as I
    lib = I.load_library()
    col = np.ascontiguousarray(inst.col, np.int32)
    rp = np.ascontiguousarray(inst.row_ptr, np.int32)
    val = np.ascontiguousarray(inst.val, np.float64)
    lhs = np.ascontiguousarray(inst.lhs, np.float64)
    rhs = np.ascontiguousarray(inst.rhs, np.float64)
    vt = np.ascontiguousarray(inst.vartype, np.int32)
    P = lambda a: ctypes.c_void_p(a.ctypes.data)  # noqa: E731
    calls = []
    for _ in range(runs):
        lb, ub = np.array(inst.lb, np.float64), np.array(inst.ub, np.float64)
        t = time.perf_counter()
        rc = lib.propagateConstraintsGPUAtomicDouble(inst.m, inst.n, len(col), P(col), P(rp), P(val), P(lhs), P(rhs),
                                                     P(lb), P(ub), P(vt), sync)
        calls.append((time.perf_counter() - t) * 1e3)
        assert rc == I.GDP_OKAY, I.last_error()
        assert np.array_equal(lb, out_lb) and np.array_equal(ub, out_ub), "drop-in call result differs"
    creates = []
    for _ in range(runs):
        t = time.perf_counter()
        p = I.Problem(inst)
        creates.append((time.perf_counter() - t) * 1e3)
        p.close()
    return {"call": f"propagateConstraintsGPUAtomicDouble(..., {SYNC_NAMES[sync]})", "call_ms": min(calls),
            "problem_create_ms": min(creates),
            "note": "host arrays in and out: layout build on the device, H2D of the CSR, propagation, D2H"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--sync", type=int, default=2, choices=[0, 1, 2])
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-runs", type=int, default=7)
    ap.add_argument("--no-cpu", action="store_true", help="skip the host baselines and the parity check")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-instances leg")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving leg")
    ap.add_argument("--no-tune", action="store_true", help="the default tuning instead of the autotuner's pick")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the other single-instance workloads (literal Zipf(1.8) shape, cover/knapsack, chains)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MB")
    ap.add_argument("--summarize", default=None, metavar="DIR")
    ap.add_argument("--summarize-batch", default=None, nargs=2, metavar=("DIR", "OUT_MD"),
                    help="write the batched leg's kernel-trace summary (no GPU)")
    ap.add_argument("--summarize-steps", default=None, nargs=2, metavar=("DIR", "OUT_MD"),
                    help="per-round step-kernel durations from a CPU_LOOP kernel trace under DIR into OUT_MD")
    ap.add_argument("--meta-out", default=None, help="write workload facts for --summarize")
    args = ap.parse_args()
    if args.summarize:
        summarize(args.summarize)
        return
    if args.summarize_steps:
        summarize_steps(*args.summarize_steps)
        return
    if args.summarize_batch:
        summarize_batch(*args.summarize_batch)
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
    tuning, default_ms = None, None
    if not args.no_tune:
        # the library's autotuner picks the layout and the persistent grid for
        # this instance (every candidate gives the same bits); the default
        # tuning's time is reported beside the tuned one
        prob = I.Problem(inst)
        dt = []
        for _ in range(4):
            lb.copy_from(root_lb)
            ub.copy_from(root_ub)
            dt.append(prob.propagate_device(lb.ptr, ub.ptr, args.sync).kernel_ms)
        default_ms = statistics.median(dt[1:])
        prob.close()
        print("bench: autotuning", file=sys.stderr, flush=True)
        t_tune = time.perf_counter()
        tuning, _ = I.autotune(inst, args.sync)
        tune_s = time.perf_counter() - t_tune
        print(f"bench: tuned in {tune_s:.2f} s: {tuning}", file=sys.stderr, flush=True)
    prob = I.Problem(inst, tuning)

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
    wall_ms = (t1 - t0) / args.steps * 1e3
    kernel = KERNEL_OF_SYNC[args.sync]
    out_lb, out_ub = lb.numpy(), ub.numpy()
    # bytes the launch has to move: each round's processed rows (row marking
    # skips the others), untimed diagnostic run of the same propagation
    print("bench: per-round work profile", file=sys.stderr, flush=True)
    wp, w_rows, w_nnz, w_cols, w_acc = prob.work_profile(inst.lb, inst.ub)
    assert wp.rounds == rounds and len(w_rows) == rounds, "work profile differs from the timed propagation"
    assert np.array_equal(wp.lbs, out_lb) and np.array_equal(wp.ubs, out_ub), "work profile result differs"
    alg = processed_bytes(w_rows, w_nnz, w_cols)
    dense = algorithmic_bytes(inst, rounds)
    achieved = alg / (kernel_ms * 1e-3) / 1e9
    print(f"bench: {rounds} rounds, kernel {kernel_ms:.3f} ms/step, wall {wall_ms:.3f} ms/step incl. bound restore, "
          f"{achieved:.1f} GB/s on processed rows ({alg / dense:.3f} of dense sweeps)", file=sys.stderr, flush=True)
    if args.meta_out:
        os.makedirs(os.path.dirname(args.meta_out) or ".", exist_ok=True)
        with open(args.meta_out, "w") as fh:
            json.dump({"workload": {"m": inst.m, "n": inst.n, "nnz": inst.nnz, "rounds": rounds},
                       "algorithmic_bytes_per_launch": alg, "dense_sweep_bytes_per_launch": dense,
                       "warmup": args.warmup, "steps": args.steps, "bench_launch_ms": kernel_ms}, fh)

    traffic = load_traffic(inst, rounds, kernel) if args.sync == 2 else None
    print("bench: device copy bandwidth", file=sys.stderr, flush=True)
    copy_gbs = measured_copy_gbs()
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
                   "sync": SYNC_NAMES[args.sync], "rounds": rounds, "global_batch": 1,
                   "tuning": tuning if tuning is not None else "default",
                   "default_tuning_ms": default_ms},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": kernel, "launch_ms": kernel_ms, "algorithmic_bytes_per_launch": alg,
                     "processed_nnz_per_launch": int(w_nnz.sum()), "processed_rows_per_launch": int(w_rows.sum()),
                     "accepted_candidates_per_launch": int(w_acc.sum()),
                     "dense_sweep_bytes_per_launch": dense,
                     "dense_equivalent_gbs": dense / (kernel_ms * 1e-3) / 1e9,
                     "measured_copy_gbs": copy_gbs, "frac_of_measured_copy": achieved / copy_gbs},
        "wall_ms_per_step": wall_ms,
    }
    print("bench: end-to-end drop-in call", file=sys.stderr, flush=True)
    result["end_to_end"] = end_to_end_leg(inst, out_lb, out_ub, args.sync)
    prob.close()
    del prob, lb, ub, root_lb, root_ub
    if not args.no_extra:
        from gdp_amd.instances import chain_mip, cover_knapsack
        extra = {}
        x = powerlaw_mip(seed=args.seed, s=1.8)
        extra["powerlaw_zipf_1_8"] = workload_leg(
            x, args.sync, 5, f"powerlaw_mip m=1e6 n=5e5 zipf_s=1.8 row_len=[2,50000] seed={args.seed} "
                             f"(the literal shape: long rows carry most nonzeros)")
        del x
        x = cover_knapsack(one_frac=0.0)
        extra["cover_knapsack"] = workload_leg(
            x, args.sync, 5, "cover_knapsack m=4e5 n=2e6, 2% of the binaries fixed to 0 (as specified it "
                             "cannot chain: 2 rounds)")
        del x
        x = chain_mip()
        extra["chains"] = workload_leg(
            x, args.sync, 5, "chain_mip m=4e5 n=2e6 row_len=[2,60] 2% pre-fixed (the many-round workload)",
            fixed_cost_rounds=8)
        del x
        result["other_workloads"] = extra
    batched_out = None
    if not args.no_batch:
        result["batched"], batched_out = batched_leg(args.sync)
    if not args.no_serve:
        result["serving"] = serving_leg(args.sync)
    if not args.no_cpu:
        parity, base, bparity = cpu_leg(inst, out_lb, out_ub, rounds, args.cpu_runs, batched_out)
        result["parity"] = parity
        result["cpu_baseline"] = base
        host = host_facts()
        result["host"] = host
        # cpu_omp on the CPUs this process may use (the GPU's share of the
        # host, at most its physical cores), and on the reference's
        # compile-time 8 threads (SHARED_MEM_THREADS)
        share = min(x for x in (host["usable_cpus"], host["physical_cores"], int(host["omp_num_threads_env"] or 0)
                                or None) if x)
        result["cpu_omp"] = omp_baseline(inst, share)
        result["cpu_omp"]["threads_are"] = "the usable CPUs of this process (the GPU's share of the host)"
        result["cpu_omp_8_threads"] = omp_baseline(inst, 8)
        if bparity is not None:
            result["batched"]["parity"] = bparity
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
