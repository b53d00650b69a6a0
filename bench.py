"""Benchmark of the ASPL/diameter evaluator on one MI355X.

Default workload (--workload random65536): one seeded random 16-regular simple
graph with n=65536, evaluated with the MATRIX method (the full 65536 x 65536 bit
matrix, 512 MiB per buffer, every column block resident).  Also: grid256 (the
256x256 4-fold symmetric grid, MATRIX_SAVING), random1024 (n=1024, d=6),
batch1024 (the batched path alone) and anneal1024 (ODPX_Anneal steps).  A step
is one full evaluation (every BFS level, exact sum, diameter and ASPL) of the
adjacency already resident in HBM.  The JSON line's value is evaluations per
second.

The timed region replays each level-plan evaluation's captured hipGraph (a
resident-plan evaluation is one direct launch).  The same steps
then run again with HIP events recorded around every launch on the library's
launch stream (eager launches, the same launch sequence: count-only, listing and
sparse last levels as they come).  roofline: the dominant kernel over those
steps (the gather level, k_level, k_sparse excluded; k_tile on grid graphs with
the tile plan, several levels per launch; k_resident_w when the graph's blocks fit
LDS); achieved = algorithmic bytes of its launches (rows * ((d+2) * k * 8 + 4 d)
per level, k = 64-bit words per row, d+1 for a count-only level; the share of
rows a frontier-mode or tile launch processed) / their summed duration.
traffic: measured HBM bytes per launch, from rocprofv3 counter runs (see
profiles/), or null.

cpu_baseline: the reference library's OpenMP build (oracle/_ref, compiled from
the reference sources) timed on this host on the same graph.

Extra objects (random65536, unless --no-extras): "batched" (1024 seeded 2-opt
candidates with n in [1024, 16384] and d in {4, 8, 12, 16, 24}, resident in HBM,
one persistent launch, against looping the single-graph GPU path) and
"served" (64 annealing chains through the serving engine: saturated rate,
then Poisson arrivals with p50/p99 latency; --no-served skips it).  Sampled
batched and served results are checked bit for bit against the reference in
the cpu_baseline leg.
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import odp_amd as odp  # noqa: E402

METRIC = "ASPL evals/s (bit-exact) at n=65536 d=16 and batched 2-opt; HBM GB/s vs peak"
HBM_PEAK_GBS = 8000.0
# expected results, from the reference library (tests/golden/vectors.json)
EXPECT = {("random", 65536, 16, 1): (6, 9346029721, float.fromhex("0x1.1689a2d622d62p+2")),
          ("random", 1024, 6, 1): (6, 2198785, float.fromhex("0x1.0cab32ccb32cdp+2")),
          ("grid", 65536, 8, 1): (128, 97430397492, float.fromhex("0x1.6af6491989199p+5"))}

WORKLOADS = {
    # name: (kind, nodes, degree, description)
    "random65536": ("random", 65536, 16, "single seeded random 16-regular simple graph, n=65536, MATRIX "
                                         "(full bit matrix, every column block resident), adjacency resident in HBM"),
    "grid256": ("grid", 65536, 8, "256x256 grid, degree 8, edge length <= 4, 4-fold rotational symmetry "
                                  "(16384 sources expanded), MATRIX_SAVING, adjacency resident in HBM"),
    "random1024": ("random", 1024, 6, "single seeded random 6-regular simple graph, n=1024 (CPU-runnable "
                                      "case), adjacency resident in HBM"),
    "anneal1024": ("anneal", 1024, 6, "simulated annealing of a seeded random 6-regular graph, n=1024, "
                                      "adjacency on the device; a step is one 2-opt move, its exact evaluation "
                                      "and the acceptance test (ODPX_Anneal)"),
    "batch1024": ("batch", 1024, 0, "1024 seeded random d-regular graphs, n uniform even in [1024,16384], "
                                    "d in {4,8,12,16,24}, one 2-opt swap each, resident in HBM; a step is one "
                                    "batched evaluation of all of them"),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cpus():
    """The host's CPU description: usable cores (affinity), nproc, lscpu's model,
    socket and core lines, and the cgroup CPU quota when one is set."""
    import subprocess
    info = {"affinity_cpus": len(os.sched_getaffinity(0)), "os_cpu_count": os.cpu_count()}
    try:
        info["nproc"] = int(subprocess.run(["nproc"], capture_output=True, text=True, timeout=10).stdout.strip())
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for ln in out.splitlines():
            k, _, v = ln.partition(":")
            if k.strip() in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core", "CPU(s)"):
                info["lscpu " + k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        quota = open("/sys/fs/cgroup/cpu.max").read().split()
        if quota[0] != "max":
            info["cgroup_cpu_quota"] = int(quota[0]) / int(quota[1])
    except (OSError, ValueError, IndexError):
        pass
    return info


def thread_candidates():
    """OpenMP thread counts to try: OMP_NUM_THREADS alone when set, else the
    usable cores, the cgroup quota and 16 (the box's share), deduplicated."""
    env = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    aff = len(os.sched_getaffinity(0))
    if env > 0 and "ODP_BENCH_THREAD_SWEEP" not in os.environ:
        return sorted({env, aff})
    cands = {aff, min(16, aff)}
    try:
        quota = open("/sys/fs/cgroup/cpu.max").read().split()
        if quota[0] != "max":
            cands.add(max(1, min(aff, int(round(int(quota[0]) / int(quota[1]))))))
    except (OSError, ValueError, IndexError):
        pass
    return sorted(cands)


def cpu_baseline(adj, gkw, kind_name, budget_s=20.0, checks=()):
    """Reference OpenMP library on a bounded sample of the same workload, at the
    best of a few thread counts (every usable core among them); also the
    reference's results for sampled graphs (`checks`: (graph, HIP result)
    pairs), compared bit for bit."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    if not oracle.ref_available(threads=True):
        return None, None
    aff = len(os.sched_getaffinity(0))
    matched = 0
    for g, res in checks:
        (d, sm, a), _ = oracle.run_reference(g, nodes=g.shape[0], threads=True, omp_threads=aff)
        if (d, sm) != tuple(res[:2]) or a != res[2]:
            raise SystemExit(f"sampled result {res} differs from the reference's {(d, sm, a)}")
        matched += 1
    kind = {"MATRIX": oracle.MATRIX, "MATRIX_SAVING": oracle.MATRIX_SAVING}[kind_name]
    sweep = {}
    cands = thread_candidates()
    per_budget = budget_s / max(1, len(cands))
    best = None
    for cores in cands:
        (d, s, a), t1 = oracle.run_reference(adj, kind=kind, threads=True, repeat=1, omp_threads=cores, **gkw)
        per = t1[0] if t1 else 1.0
        reps = max(2, min(30, int(per_budget / max(per, 1e-4))))
        (d, s, a), ts = oracle.run_reference(adj, kind=kind, threads=True, repeat=reps, omp_threads=cores, **gkw)
        med = statistics.median(ts)
        sweep[str(cores)] = {"median_s": med, "calls": reps}
        if best is None or med < best[1]:
            best = (cores, med, reps)
    cores, med, reps = best
    return ({"value": 1.0 / med, "unit": "evals/s", "cores": cores, "kind": "reference",
             "sample": f"{reps} timed ODP_Set_aspl calls (after 1 warm-up) of the reference OpenMP library, "
                       f"{kind_name}, on the same graph, at the fastest of OMP_NUM_THREADS in {cands}; "
                       f"median {med:.6f} s",
             "thread_sweep": sweep, "host": host_cpus(),
             "sampled_results_matching_reference": matched},
            (d, s, a))


def calibrate_copy(mib):
    """One hipMemcpy device-to-device of `mib` MiB (runs the runtime's copy kernel)."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    size = mib << 20
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(size)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(size)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(size)) == 0
    assert hip.hipMemcpy(b, a, ctypes.c_size_t(size), 3) == 0      # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"calibration copy: {mib} MiB device-to-device")


def c4_params(count, seed=2024):
    """(n, d) of the batched workload's graphs: n uniform even in [1024, 16384],
    d uniform in {4, 8, 12, 16, 24}."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(rng.integers(512, 8193)) * 2
        d = int(rng.choice([4, 8, 12, 16, 24]))
        out.append((n, d))
    return out


def c4_graph(g, n, d, seed=2024):
    """Graph g of the batched workload: a seeded random d-regular graph, then one 2-opt swap."""
    adj = odp.generate_regular(n, d, seed * 100000 + g)
    odp.two_opt(adj, odp.Rng(seed * 100000 + g + 7))
    return adj


def c4_graphs(count, seed=2024):
    return [c4_graph(g, n, d, seed) for g, (n, d) in enumerate(c4_params(count, seed))]


C5_NODES, C5_DEGREES = (4096, 16384, 65536), (8, 16, 32)


def c5_chain(i):
    """Start graph of served chain i: n = C5_NODES[i % 3], d = C5_DEGREES[(i // 3) % 3]."""
    return odp.generate_regular(C5_NODES[i % 3], C5_DEGREES[(i // 3) % 3], 5000 + i)


def served_extra(nchains=64, saturate_s=4.0, poisson_s=6.0, load=0.6, seed=77, small_slots=1, large_slots=2, rate=None):
    """64 annealing chains, n in {4096, 16384, 65536} x d in {8, 16, 32}, each request one
    random 2-opt swap (accepted: the next request applies to the swapped graph)."""
    t0 = time.perf_counter()
    mirrors = [c5_chain(i) for i in range(nchains)]
    log(f"served: generated {nchains} chains in {time.perf_counter() - t0:.1f}s")
    rng = odp.Rng(seed)
    srv = odp.Server(small_slots=small_slots, large_slots=large_slots, hbm_budget_gb=64)
    samples = []            # (graph copy, request id) checked against the reference later
    want_sample = {2: None, 3: None, 5: None}    # chains 2 (n=65536), 3 (n=4096), 5 (n=65536)
    try:
        cids = [srv.add_chain(m) for m in mirrors]
        sent = [0] * nchains

        def submit(i):
            sw = odp.two_opt(mirrors[i], rng)
            rid = srv.submit(cids[i], [sw])
            sent[i] += 1
            if i in want_sample and want_sample[i] is None and sent[i] == 3:
                want_sample[i] = rid
                samples.append((mirrors[i].copy(), rid))
            return rid

        large = [mirrors[i].shape[0] > 16384 for i in range(nchains)]
        n_large = sum(large)
        # closed loop: every chain keeps one request in flight
        results = {}
        for i in range(nchains):
            submit(i)
        t = time.perf_counter()
        done, done_large = 0, 0
        while time.perf_counter() - t < saturate_s:
            for r in srv.poll():
                results[r["id"]] = r
                done += 1
                done_large += large[r["chain"]]
                submit(r["chain"])
        srv.drain()
        for r in srv.poll():
            results[r["id"]] = r
        sat_rate = done / saturate_s
        x_large, x_small = done_large / saturate_s, (done - done_large) / saturate_s
        # open loop: Poisson arrivals, chains uniform; the mix saturates when its slowest
        # class does, so the rate is `load` of min over classes of capacity / class share
        cap_mix = min(x_large * nchains / max(1, n_large), x_small * nchains / max(1, nchains - n_large))
        fixed_rate = rate is not None
        rate = max(1.0, load * cap_mix) if rate is None else rate
        arr = np.random.default_rng(seed)
        lat, lat_large, starts = [], [], {}
        t = time.perf_counter()
        next_t = arr.exponential(1.0 / rate)
        submitted = 0
        while True:
            now = time.perf_counter() - t
            if now >= poisson_s:
                break
            if now >= next_t:
                rid = submit(int(arr.integers(nchains)))
                starts[rid] = True
                submitted += 1
                next_t += arr.exponential(1.0 / rate)
            for r in srv.poll():
                results[r["id"]] = r
                if r["id"] in starts:
                    (lat_large if large[r["chain"]] else lat).append(r["t_done"] - r["t_submit"])
            if next_t - (time.perf_counter() - t) > 2e-4:
                time.sleep(1e-4)
        srv.drain()
        elapsed = time.perf_counter() - t
        for r in srv.poll():
            results[r["id"]] = r
            if r["id"] in starts:
                (lat_large if large[r["chain"]] else lat).append(r["t_done"] - r["t_submit"])
        bad = sum(1 for r in results.values() if r["status"] != 0)
        lat_ms = np.array(lat + lat_large) * 1e3
        lat_s_ms = np.array(lat) * 1e3
        lat_l_ms = np.array(lat_large) * 1e3
        sample_checks = [(g, (results[rid]["diameter"], results[rid]["sum"], results[rid]["ASPL"]))
                         for g, rid in samples]
        return ({"chains": nchains, "large_chains": n_large, "saturated_evals_per_s": sat_rate,
                 "saturated_large_evals_per_s": x_large, "saturated_small_evals_per_s": x_small,
                 "poisson_rate_per_s": rate, "poisson_evals_per_s": len(lat_ms) / elapsed,
                 "p50_ms": float(np.percentile(lat_ms, 50)), "p99_ms": float(np.percentile(lat_ms, 99)),
                 "p50_ms_small": float(np.percentile(lat_s_ms, 50)), "p99_ms_small": float(np.percentile(lat_s_ms, 99)),
                 "p50_ms_large": float(np.percentile(lat_l_ms, 50)), "p99_ms_large": float(np.percentile(lat_l_ms, 99)),
                 "requests": len(results), "failed": bad, "slots": {"small": small_slots, "large": large_slots},
                 "poisson_rate_source": "fixed" if fixed_rate else "60% of the measured mix capacity",
                 "workload": "64 chains, n in {4096,16384,65536} x d in {8,16,32}, one accepted 2-opt swap per "
                             "request; closed loop for the saturated rates, then Poisson arrivals over uniform "
                             "chains"},
                sample_checks)
    finally:
        srv.stop()


def batched_extra(count, loops):
    t0 = time.perf_counter()
    graphs = c4_graphs(count)
    packed = odp.pack(graphs)
    log(f"batched: generated {count} graphs in {time.perf_counter() - t0:.1f}s")
    t = time.perf_counter()
    ng = odp.batch_load(packed)          # upload + tables: the graphs are then resident in HBM
    t_load = time.perf_counter() - t
    ref = odp.batch_run(ng)              # warm-up, and the results every other path must match

    def timed():
        times = []
        for _ in range(loops):
            t = time.perf_counter()
            r = odp.batch_run(ng)
            times.append(time.perf_counter() - t)
            assert all((x == y).all() for x, y in zip(r, ref))
        return min(times)
    odp.batch_run(ng)                    # second run: captures the launch graph
    t_default = timed()
    t = time.perf_counter()
    tried = odp.batch_autotune(2)        # launch policy per size class, checked identical
    t_tune = time.perf_counter() - t
    policy = odp.batch_policy()
    odp.batch_run(ng)
    odp.batch_run(ng)
    tb = timed()
    odp.loop_run(ng, odp.PLAN_LEVEL)     # warm-up (buffer growth)
    t = time.perf_counter()
    lp = odp.loop_run(ng, odp.PLAN_LEVEL)
    tl = time.perf_counter() - t
    odp.loop_run(ng, odp.PLAN_AUTO)
    t = time.perf_counter()
    la = odp.loop_run(ng, odp.PLAN_AUTO)
    ta = time.perf_counter() - t
    same = all((x == y).all() for x, y in zip(lp, ref)) and all((x == y).all() for x, y in zip(la, ref))
    checks = [(graphs[g], (int(ref[0][g]), int(ref[1][g]), float(ref[2][g]))) for g in (0, count // 2)]
    return checks, {"graphs": count, "evals_per_s": count / tb, "ms_per_batch": 1e3 * tb,
            "loop_level_evals_per_s": count / tl, "loop_auto_evals_per_s": count / ta,
            "speedup_vs_loop_level": tl / tb, "speedup_vs_loop_auto": ta / tb,
            "identical_results": bool(same), "load_s": t_load,
            "default_policy_ms": 1e3 * t_default, "autotune": {"candidates": tried, "seconds": t_tune,
                                                               "wb_cap_per_class": policy[0],
                                                               "workgroups_per_cu": policy[1]},
            "workload": "1024 seeded random d-regular graphs, n uniform even in [1024,16384], "
                        "d in {4,8,12,16,24}, one 2-opt swap each"}


BATCH_CLASSES = (("n<=4096", 0, 4096), ("4096<n<=10208", 4096, 10208), ("n>10208", 10208, 1 << 30))


def batch_classes():
    """The 1024-graph batch split by graph size (the 10208 cut is where blocks
    narrow from 4 to 2 words): per class the batched run against looping the
    single-graph auto plan, best of 3 (batch) and 2 (loop) runs after warm-up."""
    params, graphs = c4_params(1024), c4_graphs(1024)

    def best(f, k):
        out = []
        for _ in range(k):
            t = time.perf_counter()
            f()
            out.append(time.perf_counter() - t)
        return min(out)
    rows, tb_sum, ta_sum = [], 0.0, 0.0
    for name, lo, hi in BATCH_CLASSES:
        gs = [g for g, (n, _) in zip(graphs, params) if lo < n <= hi]
        ng = odp.batch_load(odp.pack(gs))
        ref = odp.batch_run(ng)
        odp.batch_autotune(2)
        odp.batch_run(ng)
        odp.batch_run(ng)
        tb = best(lambda: odp.batch_run(ng), 3)
        la = odp.loop_run(ng, odp.PLAN_AUTO)
        ta = best(lambda: odp.loop_run(ng, odp.PLAN_AUTO), 2)
        assert all((x == y).all() for x, y in zip(la, ref))
        odp.stats_enable(True, live=True)        # the loop's kernel time alone, from HIP events
        odp.stats_reset()
        odp.loop_run(ng, odp.PLAN_AUTO)
        st = odp.stats()
        odp.stats_enable(False)
        odp.batch_release()
        tb_sum += tb
        ta_sum += ta
        rows.append({"class": name, "graphs": len(gs), "batch_ms": round(1e3 * tb, 3),
                     "loop_auto_ms": round(1e3 * ta, 3), "loop_auto_us_per_graph": round(1e6 * ta / len(gs), 1),
                     "loop_kernel_us_per_graph": round(1e3 * st["resident_ms"] / len(gs), 1),
                     "speedup_vs_loop_auto": round(ta / tb, 3)})
        log(f"batch class {name}: {rows[-1]}")
    print(json.dumps({"batch_classes": rows, "batch_ms": round(1e3 * tb_sum, 3),
                      "loop_auto_ms": round(1e3 * ta_sum, 3), "speedup_vs_loop_auto": round(ta_sum / tb_sum, 3)}),
          flush=True)


SWEEP_NODES = (1024, 4096, 8192, 12000, 16384, 20000, 24576, 32768, 40000, 49152, 65536)
SWEEP_DEGREES = (4, 8, 16, 32)


def size_sweep(steps):
    """The library's default plan on seeded random d-regular graphs over a grid
    of sizes: the median time of `steps` evaluations of the adjacency resident
    in HBM (ODPX_Set_aspl_resident), every result equal to the first."""
    rows = []
    for n in SWEEP_NODES:
        for d in SWEEP_DEGREES:
            adj = odp.generate_regular(n, d, 11 + n + d)
            odp.init_aspl_general(n, d)
            try:
                ref = odp.set_aspl(adj)
                odp.upload_adjacency(adj)
                odp.set_aspl_resident()
                odp.set_aspl_resident()
                ts = []
                for _ in range(steps):
                    t = time.perf_counter()
                    r = odp.set_aspl_resident()
                    ts.append(time.perf_counter() - t)
                    assert r == ref, (n, d, r, ref)
                plan = odp.get_plan()
            finally:
                odp.finalize_aspl()
            rows.append({"nodes": n, "degree": d, "diameter": ref[0], "us": round(1e6 * statistics.median(ts), 1),
                         "plan": {k: v for k, v in plan.items() if v}})
            log(f"sweep n={n} d={d}: {rows[-1]['us']} us")
    print(json.dumps({"size_sweep": rows, "steps": steps}), flush=True)


def batch_workload(args, desc):
    """The batched path as the measured step (for profiling k_resident_w): one batch
    run of the 1024 graphs per step, with the tuned launch policy."""
    graphs = c4_graphs(1024)
    ng = odp.batch_load(odp.pack(graphs))
    odp.batch_autotune(1)
    ref = odp.batch_run(ng)
    for _ in range(max(2, args.warmup)):
        odp.batch_run(ng)
    t = time.perf_counter()
    for _ in range(args.steps):
        r = odp.batch_run(ng)
    el = time.perf_counter() - t
    assert all(x.tobytes() == y.tobytes() for x, y in zip(r, ref))
    # the same runs again with HIP events around the launch (eager)
    odp.stats_enable(True, live=True)
    odp.stats_reset()
    for _ in range(args.steps):
        odp.batch_run(ng)
    st = odp.stats()
    odp.stats_enable(False)
    ms = st["resident_ms"] / max(1, st["resident_launches"])
    alg = st["resident_bytes"] / max(1, st["resident_launches"])
    # compulsory HBM bytes of a launch: every graph's 16-bit neighbour table read once
    comp = float(sum(a.shape[0] * (-(-a.shape[1] // 8) * 8) * 2 for a in graphs))
    achieved = comp / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    policy = odp.batch_policy()
    odp.batch_release()
    out = {"metric": METRIC, "value": round(1024 * args.steps / el, 3), "unit": "evals/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * el / args.steps, 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
           "data": "synthetic",
           "config": {"workload": desc, "name": args.workload, "graphs": 1024,
                      "policy": {"wb_cap_per_class": policy[0], "workgroups_per_cu": policy[1]}},
           "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": None, "kernel": "k_resident_w",
                        "launch_ms": round(ms, 4), "compulsory_bytes_per_launch": comp,
                        "launches": st["resident_launches"],
                        "work_equivalent": {"GBs": round(alg / (ms * 1e-3) / 1e9, 1) if ms > 0 else 0.0,
                                            "bytes_per_launch": alg,
                                            "note": "operand bytes of next[v] = cur[v] | OR cur[u] over every "
                                                    "level, as if each were pulled; they live in LDS"},
                        "note": "the bit rows live in LDS: HBM carries each graph's ids only, so the HBM "
                                "fraction is tiny; the kernel is bound by LDS operations"},
           "cpu_baseline": None}
    print(json.dumps(out), flush=True)


def anneal_workload(args, desc, n, d):
    """The annealing loop as the measured step: args.steps moves of one chain
    (temperature 100 -> 0.2 over the run), after args.warmup moves of another."""
    adj = odp.generate_regular(n, d, args.seed)
    os.environ["ODP_ASPL"] = "MATRIX"
    odp.init_aspl_general(n, d)
    odp.anneal(adj, max(1, args.warmup), 100.0, 0.2, seed=args.seed + 1)
    t = time.perf_counter()
    best, r = odp.anneal(adj, args.steps, 100.0, 0.2, seed=args.seed)
    el = time.perf_counter() - t
    check = odp.set_aspl(best)
    odp.finalize_aspl()
    assert (check[0], check[1]) == (r["best_diameter"], r["best_sum"])
    out = {"metric": METRIC, "value": round(r["steps"] / el, 3), "unit": "annealing steps/s", "n_gpus": 1,
           "steps": r["steps"], "warmup": args.warmup, "ms_per_step": round(1e3 * el / max(1, r["steps"]), 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
           "data": "synthetic",
           "config": {"workload": desc, "name": args.workload, "nodes": n, "degree": d, "seed": args.seed,
                      "accepted": r["accepted"], "best_diameter": r["best_diameter"], "best_sum": r["best_sum"],
                      "best_aspl": r["best_ASPL"]},
           "roofline": None}
    if not args.no_cpu:
        # one evaluation is most of a step: the reference library's evaluations of the start graph
        cb, cres = cpu_baseline(adj, dict(nodes=n), "MATRIX")
        if cb is not None:
            cb["note"] = "reference evaluations of the start graph, one per annealing step"
        out["cpu_baseline"] = cb
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="random65536", choices=sorted(WORKLOADS))
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--plan", default="auto",
                    help="auto (the library's choice) | matrix | matrix-rows | "
                         "level:<kw>[:upw:seg_major:xcd:group:direct:batch:tile] | resident:<wb>:<threads>")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--no-served", action="store_true", help="skip the serving-engine workload")
    ap.add_argument("--served-slots", default="1,2", metavar="SMALL,LARGE",
                    help="serving engine slots: concurrent small-chain batches, concurrent large-chain evaluations")
    ap.add_argument("--served-rate", type=float, default=548.0,
                    help="Poisson arrivals per second for the served workload (0: 60%% of the measured capacity); "
                         "548/s is 60%% of the capacity measured with 2 small + 4 large slots, kept so that "
                         "latencies compare at one load")
    ap.add_argument("--served-only", action="store_true", help="run the served workload alone (no timed evaluation)")
    ap.add_argument("--size-sweep", action="store_true",
                    help="the default plan's time per evaluation over a grid of graph sizes and degrees")
    ap.add_argument("--batch-classes", action="store_true",
                    help="the batched workload split by graph size, each class against looping the auto plan")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "traffic.json"))
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="first run one device-to-device copy of MIB MiB (a plain copy kernel of known "
                         "size, to scale rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (the path would shard "
              f"by source-column blocks, one share per GPU, with one 16-byte reduction of the sums)", file=sys.stderr)
        sys.exit(2)

    if args.served_only:
        sm, lg = (int(x) for x in args.served_slots.split(","))
        res, _ = served_extra(small_slots=sm, large_slots=lg, rate=args.served_rate or None)
        print(json.dumps({"served": res, "slots": [sm, lg]}), flush=True)
        return
    if args.batch_classes:
        return batch_classes()
    if args.size_sweep:
        return size_sweep(args.steps)
    gkind, n, d, desc = WORKLOADS[args.workload]
    if gkind == "batch":
        return batch_workload(args, desc)
    if gkind == "anneal":
        return anneal_workload(args, desc, n, d)
    t0 = time.perf_counter()
    if gkind == "random":
        adj = odp.generate_regular(n, d, args.seed)
        gkw = dict(nodes=n)
        kind_name = "MATRIX"
        os.environ["ODP_ASPL"] = kind_name
        odp.init_aspl_general(n, d)
    else:
        full = odp.generate_grid(256, 256, d, 4, 4, args.seed)
        adj = odp.grid_base_table(256, 256, d, 4, full)
        gkw = dict(width=256, height=256, symmetries=4)
        kind_name = "MATRIX_SAVING"
        os.environ["ODP_ASPL"] = kind_name
        odp.init_aspl_grid_s(256, 256, d, None, 4)
    log(f"{args.workload}: graph generated in {time.perf_counter() - t0:.2f}s")
    if args.plan == "matrix":
        # every column block resident (the full bit matrix), 16-word slabs, XCD-owned
        odp.set_plan(plan=odp.PLAN_LEVEL, kw=16, upw=8, seg_major=0, xcd=1, group=0, direct=-1)
    elif args.plan == "matrix-rows":
        # the full bit matrix with whole 8 KiB rows (the reference's layout), every level gathered
        odp.set_plan(plan=odp.PLAN_LEVEL, kw=0, upw=4, seg_major=1, xcd=0, group=0, direct=0)
    elif args.plan.startswith("level:"):
        f = args.plan.split(":")
        odp.set_plan(plan=odp.PLAN_LEVEL, kw=int(f[1]), upw=int(f[2]) if len(f) > 2 else 8,
                     seg_major=int(f[3]) if len(f) > 3 else 0, xcd=int(f[4]) if len(f) > 4 else 1,
                     group=int(f[5]) if len(f) > 5 else 0, direct=int(f[6]) if len(f) > 6 else -1,
                     batch=int(f[7]) if len(f) > 7 else 0, tile=int(f[8]) if len(f) > 8 else 0)
    elif args.plan.startswith("resident:"):
        f = args.plan.split(":")
        odp.set_plan(plan=odp.PLAN_RESIDENT, wb=int(f[1]), threads=int(f[2]))
    plan = odp.get_plan()
    first = odp.set_aspl(adj)                      # uploads and evaluates (PCIe-inclusive)
    want = EXPECT.get((gkind, n, d, args.seed))
    if want is not None and (first[0], first[1]) != want[:2]:
        raise SystemExit(f"result {first} differs from the reference's {want}")
    log(f"first evaluation: diameter={first[0]} sum={first[1]} ASPL={first[2]!r}; plan {plan}")

    # A step is the annealing pattern: one 2-opt move applied to the adjacency in
    # HBM (ODPX_Swap_device), then one exact evaluation of the new graph -- a
    # chain of distinct neighbours.  The moves are drawn beforehand on a host
    # mirror; every timed step's result is checked against the reference later.
    # Symmetric (grid) tables take no device swaps: their step re-evaluates the graph.
    chain = gkind == "random"
    mirror = adj.copy()
    rng = odp.Rng(args.seed + 424242)
    moves = [odp.two_opt(mirror, rng) for _ in range(max(2, args.warmup))] if chain else []
    for m in moves:
        assert odp.swap_device(*m) == 0
        odp.set_aspl_resident()
    if not chain:
        for _ in range(max(2, args.warmup)):
            odp.set_aspl_resident()
    states, timed_moves = [], []
    if chain:
        for _ in range(args.steps):
            timed_moves.append(odp.two_opt(mirror, rng))
            states.append(mirror.copy())
    results = []
    t = time.perf_counter()
    for i in range(args.steps):
        if chain:
            odp.swap_device(*timed_moves[i])
        results.append(odp.set_aspl_resident())
    el = time.perf_counter() - t
    if not chain:
        assert all(r == first for r in results)
    # the same number of evaluations of one unchanged graph (the last one)
    t = time.perf_counter()
    for i in range(args.steps):
        r = odp.set_aspl_resident()
    el_same = time.perf_counter() - t
    assert r == results[-1]
    # the chain continued with HIP events around every launch on the library's
    # stream: the same launch sequence (count-only, listing and sparse last
    # levels included), launched eagerly
    live_moves = [odp.two_opt(mirror, rng) for _ in range(args.steps)] if chain else []
    odp.stats_enable(True, live=True)
    odp.stats_reset()
    t = time.perf_counter()
    for i in range(args.steps):
        if chain:
            odp.swap_device(*live_moves[i])
        odp.set_aspl_resident()
    el_live = time.perf_counter() - t
    st = odp.stats()
    odp.stats_enable(False)
    log(f"timed {args.steps} steps in {el:.3f}s ({el_same:.3f}s on one unchanged graph, {el_live:.3f}s with "
        f"kernel timing); kernel stats {st}")

    # PCIe-inclusive rate of the drop-in call (host adjacency every call), reported in config only.
    # The calls alternate between the table and the same graph with every row's neighbours
    # in reverse order, so every call sends a whole table (an array that differs from the last
    # upload in a few entries only would be sent as those entries).
    flipped = np.ascontiguousarray(adj[:, ::-1])
    odp.set_aspl(flipped)
    t = time.perf_counter()
    for i in range(10):
        r = odp.set_aspl(adj if i % 2 == 0 else flipped)
    pcie_rate = 10 / (time.perf_counter() - t)
    if (r[0], r[1]) != (first[0], first[1]):
        raise SystemExit(f"PCIe-inclusive calls gave {r}, not {first}")
    # The reference's annealing pattern through the drop-in call: a 2-opt move on
    # the caller's array, then ODP_Set_aspl on it (the library sends the changed
    # entries, not the table); the last result joins the reference checks.
    dropin_rate, dropin_check = None, None
    if chain:
        arr = adj.copy()
        rng_d = odp.Rng(args.seed + 777)
        odp.set_aspl(arr)
        t = time.perf_counter()
        for i in range(20):
            odp.two_opt(arr, rng_d)
            r_d = odp.set_aspl(arr)
        dropin_rate = 20 / (time.perf_counter() - t)
        dropin_check = (arr.copy(), r_d)
    odp.finalize_aspl()

    # The dominant kernel: the gather level (k_tile's launches of several levels
    # on grid graphs with the tile plan), or the LDS-resident kernel when the plan
    # is resident.  achieved = compulsory HBM bytes of its launches (each row of
    # a working block read once and written once, the ids once per launch) over
    # their duration; the operand-model rate is reported apart as work-equivalent.
    if st["level_launches"] > 0:
        kern = "k_tile" if plan.get("tile") else ("k_pgather" if plan.get("parts") else "k_level")
    else:
        kern = "k_resident_w"
    pre = "resident" if kern == "k_resident_w" else "level"
    nl = max(1, st[pre + "_launches"])
    launch_ms = st[pre + "_ms"] / nl
    alg = st[pre + "_bytes"] / nl
    if kern == "k_resident_w":
        wb = plan["wb"]
        blocks = -(-n // (32 * wb))
        comp = blocks * n * (-(-d // 8) * 8) * 2.0        # each block reads the 16-bit table once
    else:
        comp = st["level_min_bytes"] / nl
    achieved = comp / (launch_ms * 1e-3) / 1e9 if launch_ms > 0 else 0.0
    traffic, measured = None, None
    if os.path.exists(args.traffic):
        tr = json.load(open(args.traffic)).get(args.workload)
        if tr and tr.get("kernel") == kern:         # counters of the kernel this run measured
            traffic = tr["bytes_per_launch"]
            measured = {"bytes_per_launch": traffic, "GBs_at_this_launch_time": round(traffic / (launch_ms * 1e-3) / 1e9, 1),
                        "source": tr.get("source"),
                        "note": "rocprofv3 FETCH_SIZE + WRITE_SIZE per launch, scaled by a plain copy: bytes that "
                                "missed L2, served by the MALL or HBM (no counter separates the two)"}
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": kern, "launch_ms": round(launch_ms, 4), "compulsory_bytes_per_launch": comp,
                "launches": st[pre + "_launches"], "measured_traffic": measured,
                "work_equivalent": {"GBs": round(alg / (launch_ms * 1e-3) / 1e9, 1) if launch_ms > 0 else 0.0,
                                    "bytes_per_launch": alg,
                                    "note": "operand bytes of next[v] = cur[v] | OR cur[u] (rows * ((d+2)*k*8 "
                                            "+ 4d) per level), most of them served by L2/MALL; a work rate, "
                                            "not an HBM fraction"},
                "timed_steps_ms_per_step": round(1e3 * el_live / args.steps, 4)}
    if kern == "k_resident_w":
        roofline["note"] = ("the bit rows live in LDS: HBM carries the neighbour ids only; the kernel is bound "
                            "by LDS and launch latency")
    elif kern == "k_tile":
        roofline["note"] = "several levels per launch from LDS (temporal tiles)"

    out = {"metric": METRIC, "value": round(args.steps / el, 3), "unit": "evals/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * el / args.steps, 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
           "data": "synthetic",
           "config": {"workload": desc, "name": args.workload,
                      "step": ("one 2-opt swap on the device copy + one exact evaluation of the new graph "
                               "(a chain of distinct neighbours)") if chain
                              else "one exact evaluation of the graph resident in HBM",
                      "nodes": n, "degree": d, "seed": args.seed, "kind": kind_name, "plan": plan,
                      "diameter": first[0], "sum": first[1], "aspl": first[2],
                      "last_step_result": list(results[-1]),
                      "same_graph_evals_per_s": round(args.steps / el_same, 3),
                      "pcie_inclusive_evals_per_s": round(pcie_rate, 3),
                      "dropin_move_evals_per_s": round(dropin_rate, 3) if dropin_rate else None},
           "roofline": roofline}
    checks = []
    if chain:
        # every timed step against the reference (bit for bit), and the drop-in pattern's last call
        checks = [(g, r) for g, r in zip(states, results)] + [dropin_check]
    if not args.no_extras and args.workload == "random65536":
        bchecks, out["batched"] = batched_extra(1024, 3)
        checks += bchecks
        if not args.no_served:
            sm, lg = (int(x) for x in args.served_slots.split(","))
            out["served"], schecks = served_extra(small_slots=sm, large_slots=lg, rate=args.served_rate or None)
            checks += schecks
    if not args.no_cpu:
        cb, cres = cpu_baseline(adj, gkw, kind_name, checks=checks)
        if cb is not None:
            if cres[:2] != first[:2] or cres[2] != first[2]:
                raise SystemExit(f"reference {cres} differs from HIP {first}")
            cb["matches_gpu_result"] = True
            cb["timed_steps_matching_reference"] = len(states)
        out["cpu_baseline"] = cb
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
