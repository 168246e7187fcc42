#!/usr/bin/env python
"""Benchmark: particle histories/s of the MI355X event-based engine.

Headline workload: the continuous-energy three-layer shield (water-like
20 cm / iron-like 10 cm / concrete-like 30 cm over four synthetic nuclides,
16384-point log grids, isotropic 2 MeV point source), 1e8 histories per step,
with cross-section tables and problem data resident in HBM before timing.

Also reported on the same JSON line:
  batched_vs_plain  the 4096-slab sweep in one batched launch sequence vs one
                    problem per launch sequence (both over all 4096 problems)
  bwd_fwd           likelihood-ratio backward cost / forward cost, 24-parameter shield
  box_variants      64 design variants of the 3D box shield in one batched engine:
                    detector response and its 80-parameter gradient per variant
                    (--box-histories total, 1e9 by default: the full workload)
  xs_lookup         the standalone lookup kernel against HBM peak
  roofline          the dominant event-loop kernel: algorithmic bytes per event
                    (one event = one flight, 152 B in continuous energy) x the
                    events of a launch / its device time from per-kernel HIP
                    events; the per-particle figure of the fused kernel beside it
  tally_check       GPU tallies vs the CPU oracle's histories (the single-core
                    sample and the parallel workers' pooled; combined-sigma z scores)
  cpu_baseline      the oracle (C restatement, 1 core) on a bounded sample
  cpu_baseline_parallel
                    the same oracle as independent processes over disjoint
                    history ranges, one per host core of this process (up to
                    16), started before the GPU is used; throughput summed
"""
from __future__ import annotations

import argparse
import importlib
import json
import math
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E, AMD public spec

# Algorithmic HBM bytes per event of the event loop (an event is one flight,
# as the reference counts Ntrack, general.cpp:76): the particle state read and
# written once per event, 2 x (position 24 + direction 24 + E 8 + weight 8 +
# cell 4 + rng 8) = 152 B in continuous energy.  The fused flight kernel
# reads and writes a particle once per launch however many flights it takes,
# so its own bytes per particle are below (ALGO_BYTES["k_flight"]).
EVENT_BYTES = 152

# Algorithmic HBM bytes per item of each event-loop kernel (particle state
# read + written, queue entries); tables and geometry are cache-resident and
# not counted.
ALGO_BYTES = {
    "k_xs_lookup": 28,       # queue 4 + material 4 + E 8 in; Sigma_t 8 + grid interval 4 out
    "k_flight": 189,         # forward runs, per PARTICLE per launch: the particle's state is read once,
                             # kept in registers through up to events_per_launch flights with their
                             # collisions and lookups, and written once: queue 4, cell 4, prob 4,
                             # material 4, pos 24, dir 24, E 8, wgt 8, rng 8, count 4, Sigma_t 8,
                             # interval 4 in; pos 24, dir 24, E 8, wgt 8, rng 8, count 4, Sigma_t 8,
                             # interval 4, status 1 out
    "k_collision": 113,      # gradient runs: queue 4, cell 4, prob 4, wgt 8, rng 8, E 8, dir 24, interval 4 in;
                             # dir 24, E 8, wgt 8, rng 8, status 1 out
    "k_refill_count": 1,     # status byte per slot
    "k_refill_scatter": 5,   # status byte per slot + queue entry per live slot
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def shield_xs_dir():
    d = os.path.join(ROOT, "_bench_xs")
    decks = importlib.import_module("mc-old_amd.decks")
    if not os.path.exists(os.path.join(d, "926056.txt")):
        decks.write_shield_nuclides(d)
    return d


def cpu_baseline(mc, decks, xs, n_sample):
    from mc_testlib import oracle_run
    p = mc.Problem.from_string(decks.shield_deck(n_sample), xs)
    t = time.perf_counter()
    r = oracle_run(p, n_sample, key_offset=10 ** 12)
    dt = time.perf_counter() - t
    return r, dt


def cpu_worker(offset, n, xs, out_path):
    """One process of the parallel CPU baseline: n oracle histories of the
    shield from key offset on (no GPU use); their tally moments go to
    out_path (.npy) for the tally check."""
    from mc_testlib import oracle_run
    mc = importlib.import_module("mc-old_amd")
    decks = importlib.import_module("mc-old_amd.decks")
    p = mc.Problem.from_string(decks.shield_deck(n), xs)
    t = time.perf_counter()
    r = oracle_run(p, n, key_offset=offset)
    dt = time.perf_counter() - t
    np.save(out_path, np.stack([r.sum, r.sq]))
    print(json.dumps({"seconds": dt, "histories": n}), flush=True)


def cpu_parallel(xs, n_each):
    """The oracle as P independent processes on disjoint key ranges (P = host
    cores available to this process, at most 16), wall time around all of
    them; run before this process touches the GPU."""
    import subprocess
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    P = max(1, min(16, cores))
    import tempfile
    tmp = tempfile.mkdtemp(prefix="mc_cpu_")
    paths = [os.path.join(tmp, f"w{k}.npy") for k in range(P)]
    t = time.perf_counter()
    procs = [subprocess.Popen([sys.executable, os.path.abspath(__file__), "--cpu-worker",
                               str(2 * 10 ** 12 + k * n_each), str(n_each), xs, paths[k]],
                              stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, cwd=ROOT)
             for k in range(P)]
    outs = [pr.communicate(timeout=900) for pr in procs]
    wall = time.perf_counter() - t
    if any(pr.returncode != 0 for pr in procs):
        raise RuntimeError("cpu worker failed: " + " | ".join(o[1][-300:] for o in outs))
    # each process times its own histories (interpreter start-up excluded);
    # they run side by side, so the slowest one bounds the aggregate
    secs = [json.loads([l for l in o[0].splitlines() if l.startswith("{")][-1])["seconds"] for o in outs]
    moments = sum(np.load(p) for p in paths)
    for p in paths:
        os.remove(p)
    os.rmdir(tmp)
    return {"value": P * n_each / max(secs), "unit": "histories/s", "cores": P, "kind": "port",
            "sample": f"{P} processes x {n_each} histories of the same shield deck on disjoint keys "
                      f"(oracle C restatement), slowest {max(secs):.1f} s ({wall:.1f} s wall with start-up), "
                      f"{platform.processor() or platform.machine()}, nproc={os.cpu_count()}"}, \
        (P * n_each, moments)


def zscores(gm, gu, om, ou, min_rel=0.2):
    """Combined-sigma z scores over bins the smaller CPU sample resolves
    (relative error <= min_rel, i.e. roughly 25 or more contributions)."""
    s = np.sqrt(gu ** 2 + ou ** 2)
    ok = (s > 0) & (om > 0) & (ou <= min_rel * np.abs(om))
    return (gm[ok] - om[ok]) / s[ok]


def _pmc_pass(counter, xs, n, tmp, timeout=240):
    """One rocprofv3 counter pass over the traffic probe.  Returns
    ({kernel: summed counter}, largest copy-kernel value, probe json)."""
    import csv
    import subprocess
    d = os.path.join(tmp, counter.lower())
    cmd = ["rocprofv3", "--pmc", counter, "--kernel-trace", "--output-format", "csv", "-d", d, "-o", "run",
           "--", sys.executable, os.path.join(ROOT, "mc-old_amd", "traffic_probe.py"), str(n), xs]
    env = dict(os.environ, TMPDIR=os.environ.get("TMPDIR", "/tmp"))
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout, env=env, cwd=ROOT)
    probe = None
    for line in r.stdout.splitlines():
        if line.startswith("{"):
            probe = json.loads(line)
    path = None
    for root, _, files in os.walk(d):
        for f in files:
            if f.endswith("counter_collection.csv"):
                path = os.path.join(root, f)
    if probe is None or path is None:
        raise RuntimeError(f"{counter} pass failed (rc {r.returncode}): {r.stderr[-400:]}")
    sums, copy = {}, 0.0
    for row in csv.DictReader(open(path)):
        if row["Counter_Name"] != counter:
            continue
        name, v = row["Kernel_Name"], float(row["Counter_Value"])
        if "copyBuffer" in name:
            copy = max(copy, v)
        if "mc::" in name:
            k = name.split("mc::")[1].split("(")[0].split("<")[0]
            sums[k] = sums.get(k, 0.0) + v
    return sums, copy, probe


def measure_traffic(xs, n=10_000_000):
    """HBM bytes per particle event of every event-loop kernel from FETCH_SIZE
    and WRITE_SIZE (KiB per dispatch), each scaled by what it reads for a
    512 MiB copy.  Runs before this process touches the GPU."""
    import shutil
    import tempfile
    if shutil.which("rocprofv3") is None:
        return None
    with tempfile.TemporaryDirectory() as tmp:
        f, fc, probe = _pmc_pass("FETCH_SIZE", xs, n, tmp)
        w, wc, _ = _pmc_pass("WRITE_SIZE", xs, n, tmp)
    cf = probe["copy_bytes"] / (fc * 1024.0)     # FETCH_SIZE reads half the copy on this part
    cw = probe["copy_bytes"] / (wc * 1024.0)
    per = {}
    for k, items in probe["items"].items():
        if items and k in f:
            per[k] = {"fetch_B": f[k] * 1024.0 * cf / items, "write_B": w.get(k, 0.0) * 1024.0 * cw / items}
    return {"copy_calibration": {"fetch_factor": cf, "write_factor": cw}, "histories": n,
            "bytes_per_item": per}


def xs_lookup_bench(mc, decks, xs, n_xs=1 << 26):
    """k_xs_lookup_array over 2^26 (energy, material) pairs of the shield's
    materials: 20 algorithmic bytes per lookup (energy, material in; Sigma_t out)."""
    eng = mc.Engine(mc.Problem.from_string(decks.shield_deck(1000), xs))
    sec, _ = eng.xs_bench(n_xs, iters=20)
    gbs = 20 * n_xs / sec / 1e9
    log(f"xs lookup: {n_xs / sec:.4g} lookups/s, {gbs:.1f} GB/s ({gbs / HBM_PEAK_GBS:.1%})")
    # the random-gather ceiling: the same streams and one (or two dependent)
    # 128-byte line reads per lookup from a table of the lookup's own size
    c1, table = eng.gather_bench(n_xs, 1)
    c2, _ = eng.gather_bench(n_xs, 2)
    log(f"gather ceiling: 1 line {n_xs / c1:.4g}/s, 2 dependent lines {n_xs / c2:.4g}/s, table {table / 2**20:.2f} MiB")
    return {"lookups": n_xs, "seconds_per_launch": sec, "lookups_per_s": n_xs / sec,
            "algo_bytes_per_lookup": 20, "achieved_GBs": gbs, "frac": gbs / HBM_PEAK_GBS,
            "gather_ceiling": {"table_bytes": table,
                               "one_line_lookups_per_s": n_xs / c1, "one_line_GBs": 20 * n_xs / c1 / 1e9,
                               "two_dependent_lines_lookups_per_s": n_xs / c2,
                               "two_dependent_lines_GBs": 20 * n_xs / c2 / 1e9,
                               "lookup_frac_of_one_line_ceiling": c1 / sec,
                               "definition": "same E/material/output streams, 1 or 2 dependent 128-byte "
                                             "line reads per lookup at hashed positions of a table the "
                                             "lookup's size, no search, no arithmetic"}}


def kernel_roofline(mc, prob, H, traffic):
    """Per-kernel device time of the headline deck (HIP events around each
    kernel, a profile run of H/10 histories) and the roofline line of the
    dominant kernel; returns (roofline, per-kernel table)."""
    peng = mc.Engine(prob, profile=True)
    nprof = max(2, H // 10)
    peng.forward(n_histories=nprof)
    ks = peng.kernel_stats()
    table = {}
    for k, s in ks.items():
        if s["launches"] == 0 or s["seconds"] == 0:
            continue
        b = ALGO_BYTES.get(k)
        # the flight kernel's items are flights; its bytes are per particle, and the
        # particles it took are the lookup kernel's items (one lookup per particle)
        units = ks["k_xs_lookup"]["items"] if k == "k_flight" else s["items"]
        gbs = b * units / s["seconds"] / 1e9 if b else None
        table[k] = {"seconds": s["seconds"], "launches": s["launches"], "items": s["items"],
                    "avg_us": 1e6 * s["seconds"] / s["launches"], "algo_GBs": gbs}
        if k == "k_flight":
            table[k]["particles"] = units
    dom = max((k for k in table if k in ALGO_BYTES), key=lambda k: table[k]["seconds"])
    d = table[dom]
    epl = d["items"] / d["launches"]
    upl = d.get("particles", d["items"]) / d["launches"]
    tr = None
    if traffic and dom in traffic["bytes_per_item"]:
        t = traffic["bytes_per_item"][dom]
        tr = (t["fetch_B"] + t["write_B"]) * epl
        for k, row in table.items():
            if k in traffic["bytes_per_item"]:
                row["hbm_B_per_item"] = traffic["bytes_per_item"][k]
    avg_s = d["seconds"] / d["launches"]
    if dom == "k_flight":
        # per event (one event = one flight): EVENT_BYTES x the launch's flights
        achieved = EVENT_BYTES * epl / avg_s / 1e9
        per_particle = {"bytes_per_particle": ALGO_BYTES[dom], "particles_per_launch": upl,
                        "achieved": d["algo_GBs"], "frac": d["algo_GBs"] / HBM_PEAK_GBS,
                        "note": "the fused kernel reads and writes each particle once per launch"}
        roof = {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": tr, "bytes_per_unit": EVENT_BYTES,
                "units_per_launch": epl, "unit_of_bytes": "event (flight)", "avg_launch_us": d["avg_us"],
                "frac_per_event": achieved / HBM_PEAK_GBS, "per_particle": per_particle,
                "traffic_per_event": tr / epl if tr else None}
    else:
        roof = {"bound": "hbm", "kernel": dom, "achieved": d["algo_GBs"], "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": d["algo_GBs"] / HBM_PEAK_GBS, "traffic": tr,
                "bytes_per_unit": ALGO_BYTES[dom], "units_per_launch": upl, "unit_of_bytes": "event",
                "avg_launch_us": d["avg_us"]}
    if traffic:
        roof["traffic_source"] = ("rocprofv3 FETCH_SIZE / WRITE_SIZE over the traffic probe "
                                  f"({traffic['histories']} histories, plain launches), scaled by a "
                                  "512 MiB copy: fetch x{:.3f}, write x{:.3f}".format(
                                      traffic["copy_calibration"]["fetch_factor"],
                                      traffic["copy_calibration"]["write_factor"]))
    return roof, table


def bwd_fwd_leg(mc, decks, xs, H, use_graph):
    """Likelihood-ratio backward cost against the forward's: the
    24-parameter shield at H histories, one timed run of each after a warm-up
    at H/10, VJP weights one over all 201 tally bins."""
    p3 = mc.Problem.from_string(decks.shield_deck(H, params="24"), xs)
    e3 = mc.Engine(p3, use_graph=use_graph)
    n3 = H                                    # the differentiable-shield case runs 1e8 histories
    e3.forward(n_histories=n3 // 10)
    go = np.ones(e3.n_bins)          # VJP weights: the transmitted fraction and every leakage bin
    e3.backward(go, n_histories=n3 // 10)
    t = time.perf_counter()
    e3.forward(n_histories=n3)
    tf = time.perf_counter() - t
    t = time.perf_counter()
    b3 = e3.backward(go, n_histories=n3)
    tbw = time.perf_counter() - t
    return {"histories": n3, "forward_s": tf, "backward_s": tbw, "ratio": tbw / tf,
            "params": int(e3.n_params), "grad_out": "ones over all 201 tally bins",
            "dT_dtheta_max_rel_sigma": float(np.max(b3.grad_sigma / np.maximum(np.abs(b3.grad), 1e-300)))}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--histories", type=float, default=1e8)
    ap.add_argument("--cpu-sample", type=int, default=100000)
    ap.add_argument("--cpu-parallel-sample", type=int, default=200000,
                    help="histories per parallel CPU worker (pooled into the tally check)")
    ap.add_argument("--box-histories", type=float, default=1e9,
                    help="total histories of the 64-variant 3D box run (the full workload: 1e9)")
    ap.add_argument("--quick", action="store_true", help="headline only")
    ap.add_argument("--no-graph", action="store_true", help="plain launches instead of hipGraph replay")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 HBM traffic passes")
    ap.add_argument("--xs-only", action="store_true", help="only the standalone cross-section lookup")
    ap.add_argument("--bwd-only", action="store_true",
                    help="only the 24-parameter shield's forward and backward (for profiling the gradient run)")
    ap.add_argument("--roofline-only", action="store_true",
                    help="only the headline deck's per-kernel profile run and its roofline line "
                         "(for profiling the dominant kernel)")
    ap.add_argument("--cpu-worker", nargs=4, metavar=("OFFSET", "N", "XS", "OUT"), help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_worker:
        cpu_worker(int(args.cpu_worker[0]), int(args.cpu_worker[1]), args.cpu_worker[2], args.cpu_worker[3])
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported in this build "
              "(histories shard with no communication; see the sharding design)", file=sys.stderr)
        sys.exit(2)

    if args.xs_only or args.bwd_only:
        mc = importlib.import_module("mc-old_amd")
        decks = importlib.import_module("mc-old_amd.decks")
        xs = shield_xs_dir()
        if args.xs_only:
            print(json.dumps(xs_lookup_bench(mc, decks, xs)), flush=True)
        else:
            print(json.dumps({"bwd_fwd": bwd_fwd_leg(mc, decks, xs, int(args.histories), not args.no_graph)}),
                  flush=True)
        return
    traffic = None
    if not (args.quick or args.no_traffic):
        # child processes under rocprofv3, started before this process loads the
        # engine library or uses the GPU (the probe writes the nuclide files)
        try:
            traffic = measure_traffic(os.path.join(ROOT, "_bench_xs"))
            log(f"traffic: {json.dumps(traffic)}")
        except Exception as ex:      # the counters are a report, not a gate
            log(f"traffic passes skipped: {ex}")

    cpu_par, pooled = None, None
    if not (args.quick or args.roofline_only):
        # child processes too, so also before this process uses the GPU
        try:
            cpu_par, pooled = cpu_parallel(shield_xs_dir(), max(1000, args.cpu_parallel_sample))
            log(f"cpu oracle, {cpu_par['cores']} processes: {cpu_par['value']:.4g} hist/s")
        except Exception as ex:      # a reported baseline, not a gate
            log(f"parallel cpu baseline skipped: {ex}")

    mc = importlib.import_module("mc-old_amd")
    decks = importlib.import_module("mc-old_amd.decks")
    xs = shield_xs_dir()
    H = int(args.histories)

    # ---------------------------------------------------------- headline
    prob = mc.Problem.from_string(decks.shield_deck(H), xs)
    G = not args.no_graph
    if args.roofline_only:
        # only the per-kernel profile run, so that a profiler wrapped around
        # this process sees the dispatches the roofline line is computed from
        roof, table = kernel_roofline(mc, prob, H, None)
        print(json.dumps({"roofline": roof, "kernels": table}), flush=True)
        return
    eng = mc.Engine(prob, use_graph=G)
    for _ in range(args.warmup):
        eng.forward(n_histories=H)
    steps = []
    res = None
    t_all = time.perf_counter()
    for _ in range(args.steps):
        t = time.perf_counter()
        res = eng.forward(n_histories=H)
        steps.append(time.perf_counter() - t)
    elapsed = time.perf_counter() - t_all
    value = H * args.steps / elapsed
    log(f"shield: {H:.3g} histories/step, {np.mean(steps):.3f} s/step, {value:.4g} hist/s, "
        f"{res.flights / np.mean(steps):.4g} flights/s, T = {res.mean[0]:.5g} +- {res.uncer[0]:.2g}")
    out = {
        "metric": "particle histories/s at <=3 sigma tally agreement, plain vs batched; bwd/fwd cost",
        "value": value,
        "unit": "histories/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded synthetic nuclides, deck generated in-process)",
        "config": {"workload": "continuous-energy 3-layer shield (water 20 / iron 10 / concrete 30 cm), "
                               "4 synthetic nuclides A=1,12,16,56 on 16384-point log grids, "
                               "isotropic 2 MeV point source, transmission + 2x100-bin energy leakage",
                   "histories_per_step": H, "flights_per_step": int(res.flights),
                   "events_per_second": res.flights / np.mean(steps),
                   "bank_capacity": int(eng.info.bank_capacity)},
    }

    # ------------------------------------------------- tally agreement + CPU
    n_cpu = args.cpu_sample
    r_cpu, dt_cpu = cpu_baseline(mc, decks, xs, n_cpu)
    # the oracle's histories: this process's sample plus the parallel workers'
    # (all on keys disjoint from the GPU run's and from each other)
    n_or, s_or, q_or = n_cpu, r_cpu.sum.copy(), r_cpu.sq.copy()
    if pooled is not None:
        n_or += pooled[0]
        s_or += pooled[1][0]
        q_or += pooled[1][1]
    om = s_or / n_or
    ou = np.sqrt(np.maximum(q_or / n_or - om * om, 0.0) / (n_or - 1.0))
    z = zscores(res.mean, res.uncer, om, ou)
    out["tally_check"] = {"reference": f"oracle, {n_or} histories on keys disjoint from the GPU run's",
                          "bins": int(z.size), "bins_total": int(res.mean.size),
                          "max_abs_z": float(np.abs(z).max()) if z.size else 0.0,
                          "frac_abs_z_gt_3": float(np.mean(np.abs(z) > 3)) if z.size else 0.0,
                          "bins_resolved": "bins whose oracle estimate has relative error <= 0.2"}
    out["cpu_baseline"] = {"value": n_cpu / dt_cpu, "unit": "histories/s", "cores": 1, "kind": "port",
                           "sample": f"{n_cpu} histories of the same shield deck (oracle C restatement, "
                                     f"{dt_cpu:.1f} s, {platform.processor() or platform.machine()}, "
                                     f"nproc={os.cpu_count()})"}
    if cpu_par:
        out["cpu_baseline_parallel"] = cpu_par
    log(f"cpu oracle: {n_cpu / dt_cpu:.4g} hist/s; tally z max {out['tally_check']['max_abs_z']:.2f}")

    if not args.quick:
        # -------------------------------------------- per-kernel roofline
        out["roofline"], out["kernels"] = kernel_roofline(mc, prob, H, traffic)
        log("kernels:", json.dumps(out["kernels"]))
        # ------------------------------------------ standalone XS lookup
        out["xs_lookup"] = xs_lookup_bench(mc, decks, xs)

        # -------------------------------------- batched vs plain sweep
        L, c, N = decks.batched_slabs(4096)
        probs = [mc.Problem.from_string(decks.slab_deck(float(l), float(cc), 0, int(k)))
                 for l, cc, k in zip(L, c, N)]
        beng = mc.Engine(probs, use_graph=G)
        # the autotuner picks workgroup shape, bank capacity, drain threshold and
        # flights per particle per iteration on a 25 % sample (every problem
        # keeps its share of histories; a smaller sample favours smaller banks
        # than the full sweep wants)
        tune = beng.autotune(int(N.sum() // 4))
        tuned = {"block": tune.block, "grid_per_cu": tune.grid_per_cu, "bank_capacity": int(tune.bank_capacity),
                 "drain_divisor": tune.drain_divisor, "events_per_launch": tune.events_per_launch,
                 "sample_hist_per_s": tune.histories_per_second}
        log(f"autotune: {tuned}")
        beng.forward(n_histories=np.maximum(2, N // 1000))   # warm-up (graph capture)
        t = time.perf_counter()
        br = beng.forward()
        tb = time.perf_counter() - t
        rate_b = float(N.sum()) / tb
        sub = np.arange(4096)
        # plain: the same kernels and launch shape, one problem at a time with
        # ordinary launches (no graph replay, no batching), over all 4096 problems
        peng2 = mc.Engine([probs[i] for i in sub], use_graph=False, block=tune.block, grid_per_cu=tune.grid_per_cu,
                          bank_capacity=int(tune.bank_capacity), drain_divisor=tune.drain_divisor,
                          events_per_launch=tune.events_per_launch)
        peng2.forward(n_histories=np.maximum(2, N[sub] // 1000), mode=mc.MC_RUN_PLAIN)
        t = time.perf_counter()
        peng2.forward(mode=mc.MC_RUN_PLAIN)
        tp = time.perf_counter() - t
        rate_p = float(N[sub].sum()) / tp

        def plain_rate(**kw):
            e = mc.Engine([probs[i] for i in sub], block=tune.block, grid_per_cu=tune.grid_per_cu,
                          bank_capacity=int(tune.bank_capacity), events_per_launch=tune.events_per_launch, **kw)
            e.forward(n_histories=np.maximum(2, N[sub] // 1000), mode=mc.MC_RUN_PLAIN)
            t0 = time.perf_counter()
            e.forward(mode=mc.MC_RUN_PLAIN)
            return float(N[sub].sum()) / (time.perf_counter() - t0)
        # references: one problem at a time but each replayed as a graph, and
        # plain launches without the drain kernel (event iterations only)
        rate_pg = plain_rate(use_graph=G, drain_divisor=tune.drain_divisor)
        rate_p3 = plain_rate(use_graph=False, drain_divisor=-1)
        out["batched_vs_plain"] = {"problems": 4096, "histories": int(N.sum()), "batched_seconds": tb,
                                   "batched_hist_per_s": rate_b, "plain_problems": int(sub.size),
                                   "plain_seconds": tp, "plain_hist_per_s": rate_p, "speedup": rate_b / rate_p,
                                   "flights": int(br.flights), "autotuned": tuned,
                                   "plain_definition": "same kernels and launch shape, one problem at a time, "
                                                       "ordinary launches",
                                   "plain_graph_replay_hist_per_s": rate_pg,
                                   "speedup_vs_plain_graph_replay": rate_b / rate_pg,
                                   "plain_event_iterations_only_hist_per_s": rate_p3,
                                   "speedup_vs_event_iterations_only": rate_b / rate_p3}
        golden = os.path.join(ROOT, "tests", "golden", "batched_slabs_analytic.npz")
        if os.path.exists(golden):
            g = np.load(golden)
            ref = np.stack([g["T"], g["R"]], axis=1).ravel()
            # bins expecting fewer than 25 scoring histories have unreliable sample variance
            ok = (br.uncer > 0) & (np.repeat(N, 2) * ref >= 25)
            zz = (br.mean[ok] - ref[ok]) / br.uncer[ok]
            out["batched_vs_plain"]["tally_check_bins"] = int(ok.sum())
            out["batched_vs_plain"]["tally_check"] = {"reference": "deterministic slab solution",
                                                      "max_abs_z": float(np.abs(zz).max()),
                                                      "frac_abs_z_gt_3": float(np.mean(np.abs(zz) > 3))}
        log(f"sweep: batched {rate_b:.4g} hist/s ({tb:.2f} s), plain {rate_p:.4g} hist/s, x{rate_b / rate_p:.2f}")

        # ------------------------------------------------ backward cost
        out["bwd_fwd"] = bwd_fwd_leg(mc, decks, xs, H, G)
        bf = out["bwd_fwd"]
        log(f"bwd/fwd: {bf['backward_s']:.3f}/{bf['forward_s']:.3f} = {bf['ratio']:.2f}")

        # ------------------------------------------------ 3D box: 64 design variants batched
        nv = 64
        nbox = int(args.box_histories) // nv
        thetas = decks.box_variant_thetas(nv)
        pbs = [mc.Problem.from_string(decks.box_deck(nbox), xs) for _ in range(nv)]
        e4 = mc.Engine(pbs, use_graph=G)
        th4 = thetas.ravel()
        go4 = np.ones(e4.n_bins)
        e4.forward(n_histories=max(1, nbox // 20), theta=th4)
        e4.backward(go4, n_histories=max(1, nbox // 20), theta=th4)
        t = time.perf_counter()
        f4 = e4.forward(theta=th4)
        tf4 = time.perf_counter() - t
        t = time.perf_counter()
        b4 = e4.backward(go4, theta=th4)
        tb4 = time.perf_counter() - t
        out["box_variants"] = {"variants": nv, "histories": nv * nbox, "params_per_variant": int(e4.n_params) // nv,
                               "forward_s": tf4, "forward_hist_per_s": nv * nbox / tf4,
                               "backward_s": tb4, "backward_hist_per_s": nv * nbox / tb4, "ratio": tb4 / tf4,
                               "flights": int(f4.flights),
                               "response_rel_sigma_median": float(np.median(f4.uncer / np.maximum(f4.mean, 1e-300))),
                               "workload": "20-cell box shield, spherical detector 1 m beyond, per-(cell, nuclide) "
                                           "scale factors in [0.8, 1.2] per variant; one batched engine"}
        log(f"box: {nv} variants x {nbox} histories, forward {tf4:.3f} s, backward {tb4:.3f} s "
            f"({tb4 / tf4:.2f}x)")

    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
