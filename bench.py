#!/usr/bin/env python3
"""Benchmark: RabbitBin graph construction on one MI355X.

A step is one pass of the path over the resident synthetic assembly:
sketches (canonical 4-mer weighted ProbMinHash), chance baseline b0, the
all-pairs pass (abundance gate + register-match compare), per-row top-200,
calibration, mutual filter and edge weights.  Inputs (contig bytes, depth
features) are resident in HBM before the timed region and the edge list stays
in HBM (rbg_asm_build_graph with RBG_EDGES_ON_DEVICE).  The timed steps replay
the step's kernels from a hipGraph captured once (RBG_BUILD_CAPTURE); the same
step launched kernel by kernel, the step with the edges copied to host memory,
and the whole call from host inputs to host edges (rbg_build_graph), are
reported beside it.

Metric: PMH pairs/sec = N(N-1)/2 / step time, on the 300,000-contig workload
(1,000 genomes, lengths LogNormal(ln 6000, 1.0), 30 samples, m=500,
max-edges=200, min-edge-score 70).

Prints one JSON line.  Every kernel of the step is timed with HIP events on
the library's stream in the eager builds; `roofline` describes the kernel with
the largest time
and `kernels` every kernel above 10% of the step's kernel time, each against
its bound (HBM bytes or int8 MFMA ops per launch).
cpu_baseline times the reference's CPU path (a performance port of its fused
pass, with the oracle's other stages) on a prefix of the same workload on the
host cores, three runs, median per phase.
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
# CPUs this process may run on, read before any OpenMP library loads (with
# OMP_PROC_BIND set, OpenMP binds the initial thread to one place)
AFFINITY = len(os.sched_getaffinity(0))

PEAK_HBM_GBS = 8000.0
PEAK_INT8_TOPS = 5000.0  # dense int8 MFMA, MI355X


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def kernel_work(name, w):
    """Algorithmic work of one launch of a kernel of the step, for the
    workload statistics w: (bound, amount, unit, peak).  Bytes are the HBM
    bytes the algorithm must move (each input row read once, outputs written
    once); the MFMA work is 2*S int8 ops per prefiltered pair.  The exact
    gate's plane gathers (mostly L2 hits) are reported beside its entry."""
    n, S, W, K, NC = w["n"], w["S"], w["W"], w["K"], w["NC"]
    pairs, cand, gated, E = w["pairs"], w["candidates"], w["gated"], w["edges"]
    if name == "sketch_kernel":
        return "hbm", w["bytes"] + n * 32 * W, "GB/s", PEAK_HBM_GBS
    if name == "gate_mfma_kernel":
        return "mfma", 2.0 * S * pairs, "TOP/s", PEAK_INT8_TOPS
    if name == "gate_compare_kernel":  # candidates read and rewritten, survivors' slots, rows once
        return "hbm", cand * 16.0 + gated * 8.0 + n * (32.0 * W + 64.0 * NC + 4.0), "GB/s", PEAK_HBM_GBS
    if name == "dense_counts_kernel":  # rows once, the count matrix written
        return "hbm", n * 32.0 * W + pairs * 4.0, "GB/s", PEAK_HBM_GBS
    if name == "scatter_kernel":  # survivor read + its two list entries
        return "hbm", gated * (16.0 + 16.0), "GB/s", PEAK_HBM_GBS
    if name == "select_wave_kernel":  # lists read, top-K written (with the queued long rows)
        return "hbm", 2.0 * gated * 8 + n * K * 6.0 + n * 16.0, "GB/s", PEAK_HBM_GBS
    if name == "emit_mark_kernel":  # top-K lists read, keep bits written
        return "hbm", n * K * 6.0 + n * 8.0 * ((K + 63) // 64), "GB/s", PEAK_HBM_GBS
    if name == "emit_write_kernel":
        return "hbm", n * K * 6.0 + E * 14.0, "GB/s", PEAK_HBM_GBS
    if name == "edge_weight_pad_kernel":  # edge arrays + both rows of unit and raw depths
        return "hbm", E * 16.0 + (E + n) * 2 * 64.0 * NC, "GB/s", PEAK_HBM_GBS
    return None


def kernel_entry(name, ms, share, w, traffic):
    kw = kernel_work(name, w)
    if kw is None:
        return {"kernel": name, "ms": ms, "share": share}
    bound, amount, unit, peak = kw
    achieved = amount / (ms / 1e3) / (1e9 if unit == "GB/s" else 1e12)
    e = {"kernel": name, "ms": ms, "share": share, "bound": bound, "achieved": achieved, "peak": peak,
         "unit": unit, "frac": achieved / peak, "algorithmic": amount, "traffic": traffic}
    if name == "gate_compare_kernel":
        # what the kernel moves through the vector memory path (mostly L2 hits):
        # both plane rows of every candidate
        g = w["candidates"] * 2.0 * 32 * w["W"]
        e["gathered_bytes"] = g
        e["gathered_tb_s"] = g / (ms / 1e3) / 1e12
    return e


def profiled_traffic(workload, full_size, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary
    (FETCH_SIZE + WRITE_SIZE scaled against a device copy), when one exists
    for this workload at its full size; None otherwise."""
    path = os.path.join(HERE, "profiles", f"graph_{workload}_traffic.md")
    if not full_size or not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(f"traffic_bytes_per_launch {kernel}:"):
            return float(line.split(":")[1])
    return None


def host_cpu():
    """CPU model name and AVX-512 feature flags of the host (/proc/cpuinfo)."""
    model, flags = None, set()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name") and model is None:
                    model = line.split(":", 1)[1].strip()
                elif line.startswith("flags"):
                    flags = set(line.split(":", 1)[1].split())
                    break
    except OSError:
        pass
    return model, sorted(x for x in flags if x.startswith("avx512"))


def cpu_pipeline(orc, seqs, offsets, depth, n, threads):
    """One run of the reference's CPU path over the first n contigs: sketches,
    depth features, b0, the fused pass (performance port of the reference's
    triangle pass), edge weights.  Returns per-phase seconds and the port's
    statistics."""
    t = {}
    t0 = time.perf_counter()
    w16 = orc.fold16(orc.sketch_batch(seqs[:int(offsets[n])], offsets[:n + 1], threads=threads))
    t["sketch"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    _, unit, raw, colnorm, anynz = orc.depth_features(depth[:n])
    t["depth_features"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    b0 = orc.pmh_baseline(w16)
    t["b0"] = time.perf_counter() - t0
    S = depth.shape[1]
    t0 = time.perf_counter()
    g = orc.port_fused_graph(w16, unit, anynz, S, b0, threads=threads)
    t["graph_call"] = time.perf_counter() - t0
    st = g["stats"]
    t["pair_pass"] = st["pass_s"]
    t["calibrate"] = st["prep_s"] + st["calib_s"]
    t["mutual"] = st["mutual_s"]
    t0 = time.perf_counter()
    orc.edge_weights(g["from"], g["to"], g["sv"], unit, raw, colnorm, S, mode=0)
    t["edge_weights"] = time.perf_counter() - t0
    t["end_to_end"] = t["sketch"] + t["depth_features"] + t["b0"] + t["graph_call"] + t["edge_weights"]
    return t, st, int(len(g["from"]))


def cpu_baseline(seqs, offsets, depth, n, budget_s, threads, contigs=None):
    """The reference's CPU path timed on the host cores: a performance port
    of its fused pass (the triangle tile pass offering every pair to both
    rows' spin-locked top-K heaps, the int8/VNNI prefilter and the FP32 gate,
    the composition early exit, calibration and the mutual drain) with the
    oracle's sketch, depth-feature and edge-weight stages around it.  The
    sample is the first n_s contigs of the workload, n_s sized so that one
    end-to-end run takes about a third of budget_s; it runs three times and
    the median of each phase is reported.  `value` projects the medians to
    the full workload (pair pass by pairs, the other phases by contigs);
    `sample_value` is the sample's own rate."""
    import numpy as np
    from oracle import oracle as orc
    if contigs:
        ns = max(2500, min(n, contigs))
    else:
        pilot_n = min(n, 12000)
        tp, _, _ = cpu_pipeline(orc, seqs, offsets, depth, pilot_n, threads)
        lin = (tp["end_to_end"] - tp["pair_pass"]) / pilot_n
        quad = tp["pair_pass"] / (pilot_n * (pilot_n - 1) / 2)
        target = budget_s / 3.0
        ns = int((-lin + (lin * lin + 2 * quad * target) ** 0.5) / quad) if quad > 0 else n
        ns = max(pilot_n, min(n, ns))
    runs = []
    for r in range(3):
        runs.append(cpu_pipeline(orc, seqs, offsets, depth, ns, threads))
        log(f"[bench/cpu] run {r + 1} of 3 over {ns} contigs: {runs[-1][0]['end_to_end']:.2f} s")
    med = {k: float(np.median([r[0][k] for r in runs])) for k in runs[0][0]}
    st, edges = runs[0][1], runs[0][2]
    pairs_s = ns * (ns - 1) / 2
    pairs_n = n * (n - 1) / 2
    full_s = med["pair_pass"] * pairs_n / pairs_s + (med["end_to_end"] - med["pair_pass"]) * n / ns
    model, avx = host_cpu()
    bind, places, _ = orc.omp_info()
    return {
        "value": pairs_n / full_s,
        "unit": "pairs/s",
        "cores": threads,
        "kind": "port",
        "sample": f"first {ns} of {n} contigs of the workload, whole CPU path (sketch, depth features, b0, "
                  f"fused pass, calibration, mutual drain, edge weights) three times, median "
                  f"{med['end_to_end']:.2f} s for {pairs_s:.3e} pairs; value projects the medians to all "
                  f"{n} contigs (pair pass by pairs, other phases by contigs): {full_s:.1f} s",
        "sample_value": pairs_s / med["end_to_end"],
        "contigs": ns,
        "runs": 3,
        "phases_s": {k: round(v, 4) for k, v in med.items()},
        "projected_full_s": full_s,
        "isa": "avx512-vnni" if st["avx512"] else "scalar",
        "int8_prefilter": bool(st["q8"]),
        "pass_counts": {"int8_pruned": st["q8_pruned"], "fp32_checked": st["float_checked"],
                        "early_exit": st["ub_pruned"], "compares": st["compares"], "tile": st["tile"]},
        "edges": edges,
        "nproc": os.cpu_count(),
        "affinity": AFFINITY,
        "cpu_model": model,
        "avx512_flags": avx,
        "omp_proc_bind": {0: "false", 1: "true", 2: "primary", 3: "close", 4: "spread"}.get(bind, str(bind)),
        "omp_places": places,
    }


def pctl(xs, q):
    import numpy as np
    return float(np.percentile(np.asarray(xs, dtype=np.float64), q)) if len(xs) else None


def batched_form(ctx, steps):
    """The batched workload (64 assemblies): one batched build per step over
    all of them (autotuned schedule, replayed from a captured hipGraph) against
    the plain per-assembly path (each assembly resident, built alone), both
    with inputs already in HBM."""
    from rabbitbin_amd import api, synth
    t0 = time.perf_counter()
    asms, spec = synth.batch_workload(64)
    pairs = sum(n * (n - 1) // 2 for n, _, _ in spec)
    log(f"[bench/batched] generated {len(asms)} assemblies, {sum(n for n, _, _ in spec)} contigs in "
        f"{time.perf_counter() - t0:.1f}s")
    params = api.default_params(rand_seed=1)
    plain_ms, plain_edges = [], []
    for seqs, off, depth in asms:
        a = api.Assembly(ctx, seqs, off, depth)
        a.build_graph(params, on_device=True)
        best = None
        for _ in range(2):
            t1 = time.perf_counter()
            r = a.build_graph(params, on_device=True)
            dt = 1e3 * (time.perf_counter() - t1)
            best = dt if best is None else min(best, dt)
        plain_ms.append(best)
        plain_edges.append(int(r["n_edges"]))
        a.free()
    # host to host: the plain path's rbg_build_graph per assembly (upload,
    # build, edges back) against one batch created, built and read back
    t1 = time.perf_counter()
    for seqs, off, depth in asms:
        ctx.build_graph(seqs, off, depth, params)
    plain_host_ms = 1e3 * (time.perf_counter() - t1)
    t1 = time.perf_counter()
    bh = api.Batch(ctx, asms)
    bh.build(params)
    for k in range(len(asms)):
        bh.result(k)
    batch_host_ms = 1e3 * (time.perf_counter() - t1)
    bh.free()
    log(f"[bench/batched] host to host: plain {plain_host_ms:.0f} ms, batch {batch_host_ms:.0f} ms")
    b = api.Batch(ctx, asms)
    (tiles, rows), tune_ms = b.autotune(params, reps=2)
    first = b.build(params, capture=True)
    times = []
    for _ in range(steps):
        t1 = time.perf_counter()
        st = b.build(params, capture=True)
        times.append(1e3 * (time.perf_counter() - t1))
    assert st["replayed"] and not first["replayed"]
    assert st["edges"] == sum(plain_edges), (st["edges"], sum(plain_edges))
    b.free()
    ms = sum(times) / len(times)
    plain_total = sum(plain_ms)
    out = {"workload": "batch64: 64 assemblies, contigs log-uniform [500, 40000], 50-400 genomes, S in "
                       "{3, 10, 30}, m=500, max-edges=200",
           "assemblies": len(asms), "contigs": sum(n for n, _, _ in spec), "pairs": pairs,
           "ms_per_step": ms, "value": pairs / (ms / 1e3), "unit": "pairs/s",
           "plain_ms_total": plain_total, "plain_value": pairs / (plain_total / 1e3),
           "speedup_vs_plain": plain_total / ms, "edges": int(st["edges"]),
           "host_to_host": {"plain_ms": plain_host_ms, "batch_ms": batch_host_ms,
                            "speedup": plain_host_ms / batch_host_ms,
                            "note": "rbg_build_graph per assembly vs one batch (create, build, results), "
                                    "inputs and edge lists on the host"},
           "schedule": {"gate_tiles_per_item": tiles, "dense_rows_per_item": rows, "autotune_ms": tune_ms},
           "note": "inputs resident in HBM for both; plain = one resident assembly built at a time (best of 2)"}
    log(f"[bench/batched] {ms:.1f} ms/step vs plain {plain_total:.1f} ms: x{plain_total / ms:.2f}")
    return out


def served_form(count, rate):
    """The served workload: sessions arriving as a Poisson process, each then
    extended 5 times by 2,000 contigs (each extension submitted when the
    previous result returns).  Reports p50/p99 latency and pairs/s (pairs of
    the graphs returned: n(n-1)/2 per result) at the offered rate, the same at
    saturation (every session arriving at once), and the plain path on the
    same request sequence (every request a full rebuild from host input, one
    at a time, FIFO latencies simulated from its measured service times; a
    request's service time is the median of three builds, since single
    builds vary with the host's state)."""
    import heapq
    import threading
    import numpy as np
    from rabbitbin_amd import api, synth
    t0 = time.perf_counter()
    arrivals, pieces = synth.served_workload(count, rate=rate)
    log(f"[bench/served] generated {count} sessions ({sum(len(p[0][1]) - 1 for p in pieces)} initial contigs) "
        f"in {time.perf_counter() - t0:.1f}s")

    def run(arr):
        recs = []
        lock = threading.Lock()
        with api.Server(0) as srv:
            # warm the server's device state with one small request
            w = pieces[0][1]
            ses, r = srv.submit(*w)
            srv.wait(r)
            srv.release(ses)
            base = time.perf_counter()

            def client(k):
                d = base + arr[k] - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                ses, r = srv.submit(*pieces[k][0])
                res, info = srv.wait(r)
                with lock:
                    recs.append(info)
                for t in range(1, len(pieces[k])):
                    r = srv.extend(ses, *pieces[k][t])
                    res, info = srv.wait(r)
                    with lock:
                        recs.append(info)
                srv.release(ses)

            th = [threading.Thread(target=client, args=(k,)) for k in range(count)]
            for t in th:
                t.start()
            for t in th:
                t.join()
            st = srv.stats()
        t_first = min(i["t_submit"] for i in recs)
        t_last = max(i["t_done"] for i in recs)
        gp = sum(i["n"] * (i["n"] - 1) // 2 for i in recs)
        lat = [i["latency_ms"] for i in recs]
        big = [i["latency_ms"] for i in recs if i["n"] >= 40000 and i["n_new"] == i["n"]]
        return {"requests": len(recs), "makespan_s": t_last - t_first, "graph_pairs": gp,
                "value": gp / (t_last - t_first), "p50_ms": pctl(lat, 50), "p99_ms": pctl(lat, 99),
                "p99_ms_assemblies_ge_40k": pctl(big, 99), "steps": st["steps"],
                "mean_requests_per_step": len(recs) / max(1, st["steps"]),
                "incremental": sum(i["incremental"] for i in recs),
                "step_ms": {k: st[k] for k in ("busy_ms", "upload_ms", "build_ms", "result_ms", "save_ms")}}

    offered = run(arrivals)
    log(f"[bench/served] offered {rate}/s: {offered}")
    saturated = run(np.zeros(count))
    log(f"[bench/served] saturated: {saturated}")
    # plain path: every request a full rebuild of the session's assembly from host input
    params = api.default_params(rand_seed=1)
    service = {}
    with api.Context(0) as ctx:
        w = pieces[0][1]
        ctx.build_graph(w[0], w[1], w[2], params)  # warm
        for k in range(count):
            for t in range(len(pieces[k])):
                parts = pieces[k][:t + 1]
                seqs = np.concatenate([p[0] for p in parts])
                offs = [np.zeros(1, np.uint64)]
                base = 0
                for p in parts:
                    offs.append(p[1][1:] + np.uint64(base))
                    base += int(p[1][-1])
                offs = np.concatenate(offs)
                depth = np.concatenate([p[2] for p in parts])
                ts = []
                for _ in range(3):
                    t1 = time.perf_counter()
                    ctx.build_graph(seqs, offs, depth, params)
                    ts.append(time.perf_counter() - t1)
                service[(k, t)] = float(np.median(ts))
    T = len(pieces[0])
    ns = {(k, t): sum(len(p[1]) - 1 for p in pieces[k][:t + 1]) for k in range(count) for t in range(T)}

    def simulate(arr):
        heap = [(float(arr[k]), k, 0) for k in range(count)]
        heapq.heapify(heap)
        free, lat, last = 0.0, [], 0.0
        while heap:
            ready, k, t = heapq.heappop(heap)
            done = max(free, ready) + service[(k, t)]
            free = done
            lat.append(1e3 * (done - ready))
            last = max(last, done)
            if t + 1 < T:
                heapq.heappush(heap, (done, k, t + 1))
        return lat, last - float(min(arr))

    gp = sum(n * (n - 1) // 2 for n in ns.values())
    lat_o, span_o = simulate(arrivals)
    lat_s, span_s = simulate(np.zeros(count))
    plain = {"service_s_total": sum(service.values()), "value": gp / sum(service.values()),
             "offered": {"p50_ms": pctl(lat_o, 50), "p99_ms": pctl(lat_o, 99), "makespan_s": span_o},
             "saturated": {"p50_ms": pctl(lat_s, 50), "p99_ms": pctl(lat_s, 99), "makespan_s": span_s}}
    log(f"[bench/served] plain: {plain}")
    return {"workload": f"served: {count} sessions, Poisson arrivals {rate}/s, contigs log-uniform "
                        f"[1000, 100000], S in [3, 30], + 5 x 2000-contig extensions each",
            "unit": "pairs/s", "offered": offered, "saturated": saturated, "plain": plain,
            "speedup_vs_plain_saturated": saturated["value"] / plain["value"]}


def plain_form(args, ctx):
    """The headline line: repeated builds of the resident 300,000-contig assembly."""
    import numpy as np
    from rabbitbin_amd import api, synth

    w = synth.WORKLOADS[args.workload]
    n = args.n or w["n"]
    t0 = time.perf_counter()
    seqs, offsets, depth, _ = synth.workload(args.workload, n=n)
    log(f"[bench] generated {args.workload}: n={n} bases={len(seqs)} S={w['S']} in "
        f"{time.perf_counter() - t0:.1f}s")
    params = api.default_params(rand_seed=1)
    t0 = time.perf_counter()
    asm = api.Assembly(ctx, seqs, offsets, depth)
    log(f"[bench] upload + depth features {time.perf_counter() - t0:.2f}s")
    # Eager builds carry per-kernel HIP-event times (the roofline table); the
    # timed steps replay the build captured into a hipGraph (same kernels,
    # no launch gaps; the capture is made in the warmup).
    runs = []
    for s in range(max(args.warmup, 1)):
        r = asm.build_graph(params, on_device=True)
        runs.append(r)
        log(f"[bench] warmup {s}: {r['ms']} edges={r['n_edges']} gated={r['gated_pairs']}")
    asm.build_graph(params, on_device=True, capture=True)
    times = []
    for s in range(args.steps):
        t0 = time.perf_counter()
        r = asm.build_graph(params, on_device=True, capture=True)
        times.append(time.perf_counter() - t0)
        log(f"[bench] step {s}: {times[-1] * 1e3:.1f} ms {r['ms']}")
    assert r["n_edges"] == runs[-1]["n_edges"]
    t0 = time.perf_counter()
    r = asm.build_graph(params, on_device=True)
    ms_eager = 1e3 * (time.perf_counter() - t0)
    runs.append(r)
    log(f"[bench] eager step: {ms_eager:.1f} ms {r['ms']}")
    ms_per_step = 1e3 * sum(times) / len(times)
    pairs = n * (n - 1) // 2
    value = pairs / (ms_per_step / 1e3)
    last = runs[-1]
    # the same step with the edge list copied to host memory (PCIe-inclusive)
    t0 = time.perf_counter()
    rh = asm.build_graph(params)
    ms_host = 1e3 * (time.perf_counter() - t0)
    log(f"[bench] host-inclusive step: {ms_host:.1f} ms, edges={rh['n_edges']}")

    # per-kernel device times (HIP events around every launch), averaged over the steps
    names = [k for k, _ in runs[0]["kernels"]]
    kms = {k: sum(dict(r["kernels"])[k] for r in runs) / len(runs) for k in names}
    total_k = sum(kms.values())
    S = w["S"]
    stats = {"n": n, "S": S, "W": 16, "K": 200, "NC": (S + 15) // 16 if S > 1 else 1, "pairs": pairs,
             "candidates": int(last["candidates"]), "gated": int(last["gated_pairs"]),
             "edges": int(last["n_edges"]), "bytes": int(offsets[-1])}
    full = n == w["n"]
    kernels = [kernel_entry(k, kms[k], kms[k] / total_k, stats, profiled_traffic(args.workload, full, k))
               for k in sorted(names, key=lambda k: -kms[k]) if kms[k] >= 0.10 * total_k]
    # the largest kernel that has a roofline (kernels are sorted by time)
    dom = next((k for k in kernels if "bound" in k), None)
    roofline = None if dom is None else {
        "bound": dom["bound"], "achieved": dom["achieved"], "peak": dom["peak"], "unit": dom["unit"],
        "frac": dom["frac"], "traffic": dom["traffic"], "kernel": dom["kernel"], "kernel_ms": dom["ms"]}
    # host inputs to host edges: upload, depth features, int8 rows and
    # pair-pass order, the step, and the edge list back (rbg_build_graph)
    t0 = time.perf_counter()
    rf = ctx.build_graph(seqs, offsets, depth, params)
    ms_h2h = 1e3 * (time.perf_counter() - t0)
    assert rf["n_edges"] == last["n_edges"]
    log(f"[bench] host to host: {ms_h2h:.1f} ms")

    out = {
        "metric": "PMH pairs/sec at 300k contigs (bit-exact graph)",
        "value": value,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8/u32",
        "data": "synthetic (seeded Markov genomes, Poisson depths)",
        "config": {"workload": f"{args.workload}: {n} contigs, {w['genomes']} genomes, S={S}, m=500, "
                               f"max-edges=200, min-edge-score=70",
                   "contigs": n, "samples": S, "edges": int(last["n_edges"]),
                   "gated_pairs": int(last["gated_pairs"]), "candidates": int(last["candidates"]),
                   "stage_ms": last["ms"], "kernel_ms": kms,
                   "ms_per_step_eager": ms_eager,
                   "ms_per_step_with_edges_to_host": ms_host,
                   "ms_host_to_host": ms_h2h,
                   "note": "value: inputs resident in HBM, edges left in HBM, the step replayed from its "
                           "captured hipGraph (ms_per_step_eager: launched one kernel at a time; kernel_ms and "
                           "the roofline come from those eager builds); ms_per_step_with_edges_to_host "
                           "adds the edge list's copy to host memory; ms_host_to_host is rbg_build_graph from "
                           "host contigs and depths to host edges (upload, depth features, int8 rows, "
                           "pair-pass order, the step, edges back)"},
        "roofline": roofline,
        "kernels": kernels,
    }
    asm.free()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_child(args)
    return out


def cpu_baseline_child(args):
    """Run the CPU baseline in a child process of its own, with OpenMP threads
    spread over the cores (OMP_PROC_BIND=spread, OMP_PLACES=cores): OpenMP
    reads its placement once per process and binds the initial thread, which
    must not happen to the process that drives the GPU.  The child makes the
    same synthetic workload and touches no GPU."""
    import subprocess
    # the GPU's share of the host: 16 cores per GPU, fewer where the process
    # may run on fewer
    threads = min(16, AFFINITY)
    env = dict(os.environ, OMP_PROC_BIND="spread", OMP_PLACES="cores", OMP_NUM_THREADS=str(threads))
    cmd = [sys.executable, os.path.abspath(__file__), "--cpu-baseline-only", "--workload", args.workload,
           "--cpu-seconds", str(args.cpu_seconds), "--cpu-threads", str(threads)]
    if args.n:
        cmd += ["--n", str(args.n)]
    if args.cpu_contigs:
        cmd += ["--cpu-contigs", str(args.cpu_contigs)]
    try:
        r = subprocess.run(cmd, env=env, stdout=subprocess.PIPE, timeout=20 * args.cpu_seconds + 600)
        if r.returncode != 0:
            return {"error": f"CPU baseline exited with status {r.returncode}"}
        return json.loads(r.stdout.decode().strip().splitlines()[-1])
    except Exception as e:  # the baseline is reported, never required
        return {"error": repr(e)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="asm300k")
    ap.add_argument("--n", type=int, default=None, help="override the contig count")
    ap.add_argument("--cpu-seconds", type=float, default=30.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-contigs", type=int, default=None,
                    help="time the CPU baseline on this many contigs (default: sized from --cpu-seconds)")
    ap.add_argument("--cpu-baseline-only", action="store_true",
                    help="print the CPU baseline's JSON object and exit (no GPU use)")
    ap.add_argument("--cpu-threads", type=int, default=None)
    ap.add_argument("--forms", default="plain,batched,served",
                    help="comma list of: plain (the headline line), batched, served")
    ap.add_argument("--served-sessions", type=int, default=24)
    ap.add_argument("--served-rate", type=float, default=4.0)
    args = ap.parse_args()
    if args.cpu_baseline_only:
        from rabbitbin_amd import synth
        seqs, offsets, depth, _ = synth.workload(args.workload, n=args.n)
        threads = args.cpu_threads or min(16, AFFINITY)
        print(json.dumps(cpu_baseline(seqs, offsets, depth, len(offsets) - 1, args.cpu_seconds, threads,
                                      args.cpu_contigs)), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported by this build "
              f"(the multi-GPU sharding is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    from rabbitbin_amd import api

    ctx = api.Context(0)
    forms = args.forms.split(",")
    out = plain_form(args, ctx) if "plain" in forms else {"forms": forms}
    if "batched" in forms:
        try:
            out["batched"] = batched_form(ctx, args.steps)
        except Exception as e:  # reported beside the headline, never required for it
            out["batched"] = {"error": repr(e)}
    ctx.close()
    if "served" in forms:
        try:
            out["served"] = served_form(args.served_sessions, args.served_rate)
        except Exception as e:
            out["served"] = {"error": repr(e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
