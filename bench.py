"""Benchmark: Louvain local-move sweep on one MI355X.

Workload: seeded random geometric graph, n = 2^24 vertices in
[0,1]^2, radius sqrt(2 ln n / (pi n)), plus 1% extra uniform-random edges with U(0,1)
weights; FP64 weights, 64-bit ids; generated on the device.  Phase 0 runs with the first
threshold of threshold cycling (1e-3).

A step is one local-move sweep (distExecuteLouvainIteration over every vertex) from the
state the phase reaches after --prep iterations; the step is repeatable because the sweep
only writes targetComm / clusterWeight.  value = directed CSR entries per second over the
timed steps, inputs resident in HBM.

Also reported: the roofline of the sweep kernel (algorithmic bytes 24E + 56V per sweep at
64-bit ids, over its HIP-event-timed duration, against 8 TB/s), time to final modularity for
the whole threshold-cycling run, and the reference's own sweep on the host CPU on the same
state (oracle/_ref, OpenMP threads), with the GPU targets checked against it.

    python bench.py [--gpus 1] [--steps 20] [--warmup 3]
    python bench.py --summarize-profile DIR     (condense rocprofv3 CSV output)
    python bench.py --hbm-table KT FETCH WRITE  (per-kernel HBM GB/s from three rocprofv3 runs)
"""
import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Rows of bucket 2 (16-lane groups; degree <= 47 with the default bounds) run in the
# launch named "local_move", the dominant one on random geometric graphs; buckets 0, 1
# and 3 run in "local_move_rest" and larger degree classes in their own launches.
G16_BUCKET = 2


def g16_kernel(tuning):
    """The bucket-2 launch's kernel as rocprof names it: single graph, part 2, and whether it
    reads the per-vertex community records (the vertex_records tuning)."""
    return f"k_local_move<false, 2, {'true' if tuning.get('vertex_records') else 'false'}>"
# rocprofv3 FETCH_SIZE / WRITE_SIZE (KiB per dispatch) against the bytes a plain 1 GiB
# device copy really moves: FETCH_SIZE reads half of them, WRITE_SIZE all of them.
FETCH_SCALE = 2.0
WRITE_SCALE = 1.0
METRIC = "local-move edges/s (%HBM roofline); time to final Q ±1e-6; batched graphs/s"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cpus():
    """The host's CPUs as this process may use them: logical CPUs in the affinity mask, the
    physical cores they belong to (lscpu's CORE/SOCKET columns), the cgroup CPU quota if
    one is set, and the thread count used for the CPU baseline: the physical cores in the
    mask, capped by the quota."""
    import subprocess

    mask = sorted(os.sched_getaffinity(0))
    info = {"logical_in_mask": len(mask), "physical_in_mask": None, "cgroup_quota_cpus": None, "model": None}
    try:
        out = subprocess.run(["lscpu", "-p=CPU,CORE,SOCKET"], capture_output=True, text=True, timeout=30).stdout
        core_of = {}
        for line in out.splitlines():
            if line and not line.startswith("#"):
                cpu, core, sock = (int(x) for x in line.split(",")[:3])
                core_of[cpu] = (sock, core)
        info["physical_in_mask"] = len({core_of[c] for c in mask if c in core_of}) or None
        for line in subprocess.run(["lscpu"], capture_output=True, text=True, timeout=30).stdout.splitlines():
            if line.startswith("Model name:"):
                info["model"] = line.split(":", 1)[1].strip()
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    try:
        quota, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if quota != "max":
            info["cgroup_quota_cpus"] = int(quota) / int(period)
    except (OSError, ValueError):
        pass
    threads = info["physical_in_mask"] or len(mask)
    if info["cgroup_quota_cpus"]:
        threads = max(1, min(threads, int(info["cgroup_quota_cpus"])))
    info["threads_used"] = threads
    return info


def cpu_baseline(dg, st_state, threads, reps, gpu_target, host):
    """The reference's sweep loop on the host, same graph and state: on every usable core
    (mean of reps full sweeps) and on one thread (mean of reps sweeps over a prefix of the
    vertices holding about 1.5e8 entries, a bounded sample of the same work)."""
    from oracle import pyoracle as po
    import numpy as np

    csr = dg.download()
    g = po.Graph(csr.offs, csr.edges)
    args = (st_state["curr"], st_state["csize"], st_state["cdeg"], st_state["vdeg"], st_state["constant"])
    if po.ref_available():
        sec, tgt = po.ref_sweep_omp(g, *args, 0, g.nv, threads, reps)
        mism = int(np.count_nonzero(tgt != gpu_target))
        v1 = int(min(g.nv, max(1, np.searchsorted(g.offs, 150_000_000))))
        e1 = int(g.offs[v1])
        sec1, tgt1 = po.ref_sweep_omp(g, *args, 0, v1, 1, reps)
        mism1 = int(np.count_nonzero(tgt1[:v1] != gpu_target[:v1]))
        return {"value": g.ne / sec, "unit": "edges/s", "cores": threads, "kind": "reference",
                "sample": f"full sweep of all {g.nv} vertices ({g.ne} directed entries) from the benchmark state, "
                          f"reference distExecuteLouvainIteration (with its localCupdate atomics) under omp for "
                          f"schedule(guided), {threads} threads = physical cores in the affinity mask"
                          + (" capped by the cgroup quota" if host.get("cgroup_quota_cpus") else "")
                          + f", mean of {reps}",
                "seconds_per_sweep": sec, "target_mismatches_vs_gpu": mism, "host": host,
                "one_thread": {"value": e1 / sec1, "unit": "edges/s", "seconds": sec1, "reps": reps,
                               "sample": f"the first {v1} vertices ({e1} entries) of the same sweep",
                               "target_mismatches_vs_gpu": mism1}}
    # the reference was not built: time the single-threaded oracle on a vertex sample
    nsamp = min(g.nv, 1 << 20)
    order = np.arange(nsamp, dtype=np.int64)
    t0 = time.perf_counter()
    t, _, _, _ = po.sweep(g, *args, order=order)
    sec = time.perf_counter() - t0
    e = int(g.offs[nsamp] - g.offs[0])
    mism = int(np.count_nonzero(t[:nsamp] != gpu_target[:nsamp]))
    return {"value": e / sec, "unit": "edges/s", "cores": 1, "kind": "port",
            "sample": f"oracle sweep of the first {nsamp} vertices ({e} entries) from the benchmark state",
            "seconds_per_sweep": sec, "target_mismatches_vs_gpu": mism, "host": host}


def config1_time_to_q(ctx, args, host):
    """Time to final Q on the CPU-runnable configuration (RGG n = 65,536, threshold 1e-6, no
    coloring or early termination), on this host: the reference's own main() (oracle/_ref,
    1 MPI rank; its "Average time for clustering", main.cpp:522-528) at one thread and on
    every usable core, against the device's whole run from the resident graph."""
    import tempfile

    from oracle import pyoracle as po

    n = 65536
    dg = ctx.gen_rgg(args.seed, n, 0.0)
    dg.run()  # warm the kernels and the allocator
    t0 = time.perf_counter()
    r = dg.run()
    dev_s = time.perf_counter() - t0
    out = {"workload": f"RGG n={n} r=sqrt(2 ln n/(pi n)), seed {args.seed}, threshold 1e-6",
           "device": {"seconds": dev_s, "final_q": r.final_q, "calls": len(r.call_q),
                      "total_iters": r.total_iters}}
    if po.ref_available():
        csr = dg.download()
        with tempfile.TemporaryDirectory() as d:
            path = os.path.join(d, "rgg65536.bin")
            po.write_bin(path, po.Graph(csr.offs, csr.edges))
            for label, thr in (("one_thread", 1), ("all_cores", host["threads_used"])):
                secs = []
                for _ in range(3):
                    recs, text = po.ref_trace_run(path, [], threads=thr, timeout=600)
                    m = re.search(r"Average time for clustering \(secs\.\): (\S+)", text)
                    secs.append(float(m.group(1)))
                q = float([x for x in recs if x["kind"] != "rebuild"][-1]["q"])
                out[label] = {"threads": thr, "clustering_seconds": sum(secs) / len(secs), "reps": len(secs),
                              "final_q": q, "abs_diff_final_q_vs_device": abs(q - r.final_q)}
            out["speedup_vs_all_cores"] = out["all_cores"]["clustering_seconds"] / dev_s
    dg.free()
    return out


def iteration_timing(dg, st, V, E, steps=3):
    """Whole Louvain iterations (sweep, the per-community delta folds that replace
    localCupdate / distUpdateLocalCinfo, modularity, rotation) continued from the benchmark
    state, back to back as in a run, each timed alone on the host; each iteration's movers come
    from the library's per-iteration record afterwards (a download of the state between
    iterations would evict the Infinity Cache a run keeps warm).  Algorithmic bytes: 24E + 56V
    (sweep) + 8V (mover flags) + 64 B per mover (two 32-byte events) + 24V (modularity)."""
    its = []
    st.sweep()  # untimed, changes nothing: the caches as the previous iteration leaves them
    for _ in range(steps):
        t0 = time.perf_counter()
        q, done = st.step()
        sec = time.perf_counter() - t0
        its.append({"ms": 1e3 * sec, "q": q})
        if done:
            break
    _, moves = dg.last_trace()
    for it, movers in zip(its, moves[-len(its):]):
        b = 24.0 * E + 88.0 * V + 64.0 * int(movers)
        it.update(movers=int(movers), bytes=b, hbm_frac=b / (it["ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS)
    ms = sum(i["ms"] for i in its) / len(its)
    return {"ms_per_iteration": ms, "edges_per_s": E / (ms * 1e-3),
            "hbm_frac": sum(i["hbm_frac"] for i in its) / len(its), "iterations": its,
            "bytes_formula": "24E + 88V + 64 per mover over the iteration's wall time"}


LAST_N = 5  # timed sweeps in a profiling run (bench.py --steps 5)


def summarize_profile(d):
    """Condenses rocprofv3 CSV output under d (kernel stats, kernel trace, counter
    collection) into a short text summary: per kernel, dispatches and mean duration, and
    per (kernel, counter) the mean value per dispatch."""
    import collections
    import csv
    import glob

    def short(name):
        name = name.replace("(anonymous namespace)::", "")
        name = name.split("(")[0] if not name.startswith("void rocprim") else name[:60]
        return name[:90]

    lines = []
    for f in sorted(glob.glob(os.path.join(d, "**", "*.csv"), recursive=True)):
        base = os.path.basename(f)
        if base.endswith("kernel_stats.csv"):
            lines.append(f"## {base}: kernel | calls | mean ms | total ms | % of kernel time")
            with open(f) as fh:
                for i, row in enumerate(csv.DictReader(fh)):
                    if i < 25:
                        lines.append(f"{short(row['Name'])} | {row['Calls']} | {float(row['AverageNs']) / 1e6:.4f} | "
                                     f"{float(row['TotalDurationNs']) / 1e6:.3f} | {float(row['Percentage']):.2f}")
        elif base.endswith("kernel_trace.csv"):
            # the local-move launches one by one, in dispatch order: the timed sweeps of a
            # bench run are the last ones (profile with --no-iter --no-ttq)
            per = collections.defaultdict(list)
            with open(f) as fh:
                rows = [r for r in csv.DictReader(fh) if "local_move" in r.get("Kernel_Name", "")]
            rows.sort(key=lambda r: int(r.get("Start_Timestamp", 0) or 0))
            for r in rows:
                per[short(r["Kernel_Name"])].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e6)
            if per:
                lines.append(f"## {base}: local-move launches: kernel | dispatches | last {LAST_N} ms each | "
                             f"mean of the last {LAST_N}")
                for k, v in per.items():
                    tail = v[-LAST_N:]
                    lines.append(f"{k} | {len(v)} | {' '.join(f'{x:.4f}' for x in tail)} | "
                                 f"{sum(tail) / len(tail):.4f}")
        elif base.endswith("counter_collection.csv"):
            agg = collections.defaultdict(lambda: collections.defaultdict(list))
            with open(f) as fh:
                rows = list(csv.DictReader(fh))
            if rows and "Dispatch_Id" in rows[0]:
                rows.sort(key=lambda r: int(r["Dispatch_Id"] or 0))
            for row in rows:
                agg[short(row.get("Kernel_Name", "?"))][row.get("Counter_Name", "?")].append(
                    float(row.get("Counter_Value", 0)))
            lines.append(f"## {base}: kernel | counter | dispatches | mean per dispatch")
            for k, cs in agg.items():
                for c, v in cs.items():
                    lines.append(f"{k} | {c} | {len(v)} | {sum(v) / len(v):.6g}")
            # the local-move launches again, over their last dispatches only (the timed
            # sweeps); listed after the means, so a reader of these lines takes these
            last = [(k, c, v[-LAST_N:]) for k, cs in agg.items() if "local_move" in k for c, v in cs.items()]
            if last:
                lines.append(f"## {base}, local-move launches, last {LAST_N} dispatches: kernel | counter | "
                             f"dispatches | mean per dispatch")
                for k, c, v in last:
                    lines.append(f"{k} | {c} | {len(v)} | {sum(v) / len(v):.6g}")
    return "\n".join(lines) + "\n"


def _short_kernel(name):
    name = name.replace("(anonymous namespace)::", "")
    return (name.split("(")[0] if not name.startswith("void rocprim") else name[:60])[:90]


def hbm_table(kt_dir, fetch_dir, write_dir, top=30):
    """Achieved HBM rate of every kernel of one program: dispatches and mean duration from a
    rocprofv3 --kernel-trace --stats run, HBM bytes per dispatch from FETCH_SIZE x FETCH_SCALE
    + WRITE_SIZE x WRITE_SCALE (KiB, mean per dispatch) from two --pmc runs of the same
    program, joined on the full kernel name; GB/s = bytes / mean duration, against 8 TB/s."""
    import collections
    import csv
    import glob

    def csvs(d, suffix):
        return sorted(glob.glob(os.path.join(d, "**", "*" + suffix), recursive=True))

    stats = {}
    for f in csvs(kt_dir, "kernel_stats.csv"):
        for row in csv.DictReader(open(f)):
            stats[row["Name"]] = (int(row["Calls"]), float(row["AverageNs"]))

    def counter(d, cname):
        agg = collections.defaultdict(list)
        for f in csvs(d, "counter_collection.csv"):
            for row in csv.DictReader(open(f)):
                if row.get("Counter_Name") == cname:
                    agg[row["Kernel_Name"]].append(float(row["Counter_Value"]))
        return {k: sum(v) / len(v) for k, v in agg.items()}

    fetch, write = counter(fetch_dir, "FETCH_SIZE"), counter(write_dir, "WRITE_SIZE")
    rows = []
    for name, (calls, ns) in stats.items():
        if name in fetch and name in write and ns > 0:
            b = 1024.0 * (fetch[name] * FETCH_SCALE + write[name] * WRITE_SCALE)
            rows.append((calls * ns, name, calls, ns, b))
    rows.sort(reverse=True)
    lines = ["## kernel | calls | mean ms | HBM MB per dispatch | GB/s | % of 8 TB/s"]
    for _, name, calls, ns, b in rows[:top]:
        gbs = b / ns
        lines.append(f"{_short_kernel(name)} | {calls} | {ns / 1e6:.4f} | {b / 1e6:.2f} | {gbs:.1f} | "
                     f"{100 * gbs / HBM_PEAK_GBS:.1f}")
    return "\n".join(lines) + "\n"


def reference_run(kind, seed, n, param, flags):
    """The reference main() run on a benchmark graph, when the committed vectors cover this
    exact input (generator, seed, n, extra-edge percentage for RGG or m for BA, flags):
    per-call Q, iterations and wall time."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden", "reference_large.json")
    try:
        cases = json.load(open(path))["cases"]
    except (OSError, ValueError, KeyError):
        return None
    for case in cases.values():
        sp = case["spec"]
        ckind = sp.get("kind", "rgg")
        cparam = sp.get("m") if ckind == "ba" else sp.get("extra_pct")
        if (ckind, sp["seed"], sp["n"], cparam, sp["flags"]) == (kind, seed, n, param, list(flags)):
            return {"q": [c["q"] for c in case["calls"]], "iters": [c["iters"] for c in case["calls"]],
                    "seconds": case.get("reference_seconds"), "threads": sp["threads"]}
    return None


def compare_reference(r, ref, source):
    """A device run's per-call Q and iterations against the reference's on the same graph."""
    dq = max(abs(a - b) for a, b in zip(r.call_q, ref["q"])) if len(ref["q"]) == len(r.call_q) else None
    return {"final_q": ref["q"][-1], "abs_diff_final_q": abs(r.final_q - ref["q"][-1]), "max_abs_diff_call_q": dq,
            "same_iterations": ref["iters"] == r.call_iters, "build_container_seconds": ref["seconds"],
            "source": source}


def batched_form(ctx, args):
    """The batched form: args.batch_graphs seeded RGGs, n log-uniform in [2^8, 2^16], every
    other graph on average with 1% extra edges; all run through the threshold-cycling loop
    at once (vt_batch_run), then a strided sample of them one graph at a time (vt_run), whose
    results must be identical."""
    import numpy as np
    import vite_amd

    rng = np.random.default_rng(args.seed)
    count = args.batch_graphs
    ns = np.rint(2.0 ** rng.uniform(8.0, 16.0, size=count)).astype(np.int64)
    extras = np.where(rng.random(count) < 0.5, 1.0, 0.0)
    t0 = time.perf_counter()
    graphs = [ctx.gen_rgg(args.seed * 1000003 + k, int(ns[k]), float(extras[k])) for k in range(count)]
    gen_s = time.perf_counter() - t0
    b = vite_amd.Batch(ctx, graphs)
    log(f"batch: {count} graphs, {b.total_nv} vertices, {b.total_ne} directed entries (generated in {gen_s:.1f}s)")
    saved = ctx.tuning
    tuning = saved
    if not args.no_autotune:
        t0 = time.perf_counter()
        tuning = b.autotune(prep=1, reps=2)
        log(f"batch autotune: {tuning} in {time.perf_counter() - t0:.2f}s")
    kw = dict(threshold_cycling=True)
    t0 = time.perf_counter()
    res, st = b.run(**kw)
    batch_s = time.perf_counter() - t0
    log(f"batch: {batch_s:.3f}s, {count / batch_s:.1f} graphs/s, {st['phases']} phases, {st['rounds']} rounds")
    stride = max(1, count // max(1, args.batch_sample))
    sample = list(range(0, count, stride))[:args.batch_sample]
    t0 = time.perf_counter()
    alone = [graphs[k].run(**kw) for k in sample]
    plain_s = time.perf_counter() - t0
    same = all(a.call_q == res[k].call_q and a.call_iters == res[k].call_iters and a.final_q == res[k].final_q
               for a, k in zip(alone, sample))
    plain_gps = len(sample) / plain_s
    log(f"one graph per launch: {len(sample)} graphs in {plain_s:.3f}s, {plain_gps:.1f} graphs/s; identical={same}")
    b.free()
    for g in graphs:
        g.free()
    ctx.set_tuning(**saved)
    return {"graphs": count, "graphs_per_s": count / batch_s, "seconds": batch_s,
            "library_seconds": st["seconds"], "library_graphs_per_s": count / st["seconds"], "tuning": tuning,
            "one_per_launch_graphs_per_s": plain_gps, "one_per_launch_sample": len(sample),
            "speedup": (count / batch_s) / plain_gps, "identical_per_graph_results": same,
            "phases": st["phases"], "rounds": st["rounds"], "total_nv": b.total_nv, "total_ne": b.total_ne,
            "workload": "RGG n log-uniform in [2^8, 2^16], 1% extra edges on half, threshold cycling"}


def skewed_form(ctx, args):
    """Skewed degrees: seeded Barabasi-Albert graph, n = args.skewed_n, m = 8, unit weights.
    Local-move sweeps (plain phase 0 after --prep iterations) with the per-kernel split over
    the degree classes, then the whole loop with coloring (8 colors) and early termination
    with probability (-t 2 -a 0.25)."""
    import numpy as np

    t0 = time.perf_counter()
    dg = ctx.gen_ba(args.seed + 3, args.skewed_n, 8)
    gen_s = time.perf_counter() - t0
    deg = np.diff(dg.offsets())
    log(f"skewed: BA n={dg.nv} m=8, {dg.ne} entries, max degree {int(deg.max())} (generated in {gen_s:.2f}s)")
    saved = ctx.tuning
    tuning = saved
    if args.skewed_tuning:
        ctx.set_tuning(**args.skewed_tuning)
        tuning = ctx.tuning
        log(f"skewed tuning as given: {tuning}")
    elif not args.no_autotune:
        t0 = time.perf_counter()
        tuning = ctx.autotune(dg, prep=args.prep, reps=3)
        log(f"skewed autotune: {tuning} in {time.perf_counter() - t0:.2f}s")
    st = dg.phase_begin(lower=-1.0, thresh=1e-3)
    for _ in range(args.prep):
        st.step()
    st.sweep()
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        st.sweep()
    wall = (time.perf_counter() - t0) / reps
    # the per-kernel split from five more sweeps of the same state, timed launch by launch: the
    # event pair around each of a sweep's ~30 launches adds time of its own, so the wall time
    # above is taken without them
    ctx.set_profiling(True)
    ctx.reset_stats()
    for _ in range(reps):
        st.sweep()
    kernels = {}
    for k in ("local_move", "local_move_rest", "local_move_mid", "local_move_wide", "local_move_hub_sort",
              "local_move_hub"):
        avg, cnt, _ = ctx.kernel_stats(k)
        if cnt:
            kernels[k] = avg
    ctx.set_profiling(False)
    # the sweep's reads at random places: one community read per entry (curr[tail]) and one
    # community-degree read per distinct neighbour community of each vertex (K of them)
    curr = st.state()["curr"]
    st.end(want_cvect=False)
    tails = dg.download().edges["tail"]
    heads = np.repeat(np.arange(len(deg), dtype=np.int64), deg)
    K = int(len(np.unique(heads * len(deg) + curr[tails])))
    del heads, tails
    r, run_s = None, None
    if not args.no_skewed_run:
        t0 = time.perf_counter()
        r = dg.run(coloring=True, max_colors=8, et_type=2, et_delta=0.25)
        run_s = time.perf_counter() - t0
        log(f"skewed: -c 8 -t 2 -a 0.25 run {run_s:.2f}s Q={r.final_q:.9f}")
    dg.free()
    ctx.set_tuning(**saved)
    log(f"skewed: sweep {1e3 * wall:.3f} ms ({dg.ne / wall / 1e9:.2f} G edges/s) {kernels}")
    # the same algorithmic bytes as the headline (24E + 56V) over the whole sweep's wall time
    sweep_bytes = 24.0 * float(deg.sum()) + 56.0 * float(len(deg))
    out = {"n": int(len(deg)), "directed_entries": int(deg.sum()), "max_degree": int(deg.max()),
           "sweep_ms": 1e3 * wall, "sweep_edges_per_s": float(deg.sum()) / wall,
           "sweep_hbm_frac": sweep_bytes / wall / 1e9 / HBM_PEAK_GBS, "kernel_avg_ms": kernels,
           "candidate_communities": K, "random_reads": int(deg.sum()) + K,
           "random_reads_per_s": (float(deg.sum()) + K) / wall,
           "flags": "-c 8 -t 2 -a 0.25", "tuning": tuning}
    if r is not None:
        out.update(run_seconds=run_s, final_q=r.final_q, levels=len(r.level_q), colors=r.num_colors)
    ref = reference_run("ba", args.seed + 3, args.skewed_n, 8, ["-c", "8", "-t", "2", "-a", "0.25"])
    if ref and r is not None:
        out["reference"] = compare_reference(
            r, ref, "the reference's main() -c 8 -t 2 -a 0.25 on the same graph (written to .bin), 1 MPI rank, "
                    f"{ref['threads']} OpenMP thread, run once in the build container")
    return out


def serving_form(args):
    """The serving form: a Poisson stream of requests (seeded RGG or BA, 50/50, n log-uniform
    in [2^10, 2^20], threshold cycling) against a server with args.serve_slots slots and a
    200 GB budget, which also batches queued small requests (n <= 2^16), and the same stream
    against the sequential solver (one slot, one request at a time).  The offered load is
    args.serve_load times the sequential solver's capacity, measured first on the same
    requests back to back."""
    import numpy as np
    import vite_amd

    rng = np.random.default_rng(args.seed + 5)
    count = args.serve_requests
    reqs = []
    for k in range(count):
        n = int(round(2.0 ** rng.uniform(10.0, 20.0)))
        reqs.append(("rgg", 7000 + k, n, 0.0) if rng.random() < 0.5 else ("ba", 9000 + k, n, 8))
    kw = dict(threshold_cycling=True)
    budget = 200 * 10**9

    def replay(slots, arrivals):
        # one slot: the sequential solver, one request at a time; more: the serving loop,
        # which also batches queued small requests
        with vite_amd.Server(hbm_budget=budget, slots=slots, max_batch=1 if slots == 1 else 32) as srv:
            t0 = time.perf_counter()
            tickets = []
            for r, at in zip(reqs, arrivals):
                while True:
                    dt = at - (time.perf_counter() - t0)
                    if dt <= 0:
                        break
                    time.sleep(min(dt, 0.01))
                tickets.append(srv.submit(*r, **kw))
            out = [srv.wait(t) for t in tickets]
        lat = np.array([i["end_s"] - i["submit_s"] for _, i in out])
        span = max(i["end_s"] for _, i in out) - min(i["submit_s"] for _, i in out)
        batched = [i["batch"] for _, i in out if i["batch"] > 1]
        return out, {"p50_ms": 1e3 * float(np.percentile(lat, 50)), "p99_ms": 1e3 * float(np.percentile(lat, 99)),
                     "graphs_per_s": count / span, "makespan_s": span, "requests_in_batched_jobs": len(batched),
                     "mean_batch": float(np.mean(batched)) if batched else 1.0}

    # capacities: the requests all at once, one slot (the sequential solver) and all slots
    _, closed = replay(1, [0.0] * count)
    _, closed_mix = replay(args.serve_slots, [0.0] * count)
    lam = args.serve_load * closed["graphs_per_s"]
    arrivals = np.cumsum(rng.exponential(1.0 / lam, size=count))
    seq_out, seq = replay(1, arrivals)
    mix_out, mix = replay(args.serve_slots, arrivals)
    same = all(a.call_q == b.call_q and a.call_iters == b.call_iters and a.final_q == b.final_q
               for (a, _), (b, _) in zip(seq_out, mix_out))
    ok = all(i["status"] == 0 for _, i in mix_out)
    log(f"serving: capacity {closed['graphs_per_s']:.1f} graphs/s sequential, {closed_mix['graphs_per_s']:.1f} with "
        f"{args.serve_slots} slots ({closed_mix['requests_in_batched_jobs']} requests in batched jobs, mean batch "
        f"{closed_mix['mean_batch']:.1f})")
    log(f"serving: offered {lam:.1f} req/s; sequential p50 {seq['p50_ms']:.1f} ms p99 {seq['p99_ms']:.1f} ms "
        f"{seq['graphs_per_s']:.1f} graphs/s; {args.serve_slots} slots p50 {mix['p50_ms']:.1f} ms "
        f"p99 {mix['p99_ms']:.1f} ms {mix['graphs_per_s']:.1f} graphs/s; identical={same}")
    return {"requests": count, "offered_per_s": lam, "slots": args.serve_slots, "budget_bytes": budget,
            "p50_ms": mix["p50_ms"], "p99_ms": mix["p99_ms"], "graphs_per_s": mix["graphs_per_s"],
            "sequential": seq, "sequential_capacity_per_s": closed["graphs_per_s"],
            "capacity_per_s": closed_mix["graphs_per_s"], "capacity_run": closed_mix,
            "identical_alone_vs_mix": same, "all_ok": ok,
            "workload": "Poisson arrivals; RGG or BA (m=8) 50/50, n log-uniform in [2^10, 2^20]; threshold cycling"}


def degree_buckets(deg, bucket_max_degree):
    """Bucket of each row as the sweep assigns it: the first bucket b with degree <=
    bucket_max_degree[b] (-1 leaves b empty), else 6 (mid and wider)."""
    import numpy as np
    b = np.full(len(deg), 6, dtype=np.int8)
    for i in range(5, -1, -1):
        if bucket_max_degree[i] >= 0:
            b[deg <= bucket_max_degree[i]] = i
    return b


def measured_traffic(kernel, workload_tag, symbol):
    """HBM bytes per launch of the kernel `symbol` (as rocprofv3 names it, without "vt::")
    from the committed rocprof counter summary of launch `kernel` on this workload
    (FETCH_SIZE x FETCH_SCALE + WRITE_SIZE x WRITE_SCALE), or None."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload_tag}_fetch_write.txt")
    if not os.path.exists(path):
        return None, None
    vals = {}
    for line in open(path):
        parts = [p.strip() for p in line.split("|")]
        if len(parts) == 4 and parts[0].endswith("vt::" + symbol) and parts[1] in ("FETCH_SIZE", "WRITE_SIZE"):
            vals[parts[1]] = float(parts[3])
    if len(vals) != 2:
        return None, None
    return 1024.0 * (vals["FETCH_SIZE"] * FETCH_SCALE + vals["WRITE_SIZE"] * WRITE_SCALE), os.path.relpath(path, ROOT)


def gather_probe(args):
    """The random-read floor of the skewed run: one 4-byte read of a per-vertex array at
    every CSR entry's tail (curr[tail], as the local move reads each neighbour's community),
    done by a plain device gather (torch.index_select) with nothing else in the kernel.  The
    same reads with the vertices renumbered by degree (hubs first) and at random show how
    much vertex order matters.  Mean of 10 launches each, timed with device events."""
    import subprocess
    import tempfile

    import numpy as np
    import torch

    # the graph comes from the library's device generator in a child process, so that this
    # process touches the GPU through torch only
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "ba.npz")
        code = ("import sys, numpy as np; sys.path.insert(0, %r); import vite_amd; "
                "ctx = vite_amd.Context(0); g = ctx.gen_ba(%d, %d, 8).download(); "
                "np.savez(%r, offs=g.offs, tails=g.edges['tail'])") % (
                    os.path.dirname(os.path.abspath(__file__)), args.seed + 3, args.skewed_n, path)
        subprocess.run([sys.executable, "-c", code], check=True)
        z = np.load(path)
        offs, tails = z["offs"], z["tails"].astype(np.int64)
    n, ne = len(offs) - 1, len(tails)
    deg = np.diff(offs)
    rank = np.empty(n, np.int64)
    rank[np.argsort(-deg, kind="stable")] = np.arange(n)
    perm = np.random.default_rng(args.seed).permutation(n)
    curr = torch.arange(n, dtype=torch.int32, device="cuda")
    out = {"probe": "gather", "graph": f"BA n={n} m=8 seed {args.seed + 3}", "reads": ne}
    for label, t in (("as_generated", tails), ("degree_ranked", rank[tails]), ("random_order", perm[tails])):
        idx = torch.from_numpy(np.ascontiguousarray(t)).to("cuda")
        torch.index_select(curr, 0, idx)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(10):
            torch.index_select(curr, 0, idx)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / 10
        out[label] = {"ms": ms, "reads_per_s": ne / (ms * 1e-3)}
        log(f"gather probe {label}: {ms:.3f} ms, {ne / ms / 1e6:.1f} G reads/s")
    return out


def main():
    if len(sys.argv) == 3 and sys.argv[1] == "--summarize-profile":
        sys.stdout.write(summarize_profile(sys.argv[2]))
        return
    if len(sys.argv) == 5 and sys.argv[1] == "--hbm-table":
        sys.stdout.write(hbm_table(*sys.argv[2:5]))
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1 << 24)
    ap.add_argument("--extra", type=float, default=1.0)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--prep", type=int, default=2, help="phase iterations before the timed sweeps")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="0 = the physical cores in the affinity mask, capped by the cgroup CPU quota")
    ap.add_argument("--cpu-reps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ttq", action="store_true", help="skip the time-to-final-Q run")
    ap.add_argument("--no-iter", action="store_true", help="skip the whole-iteration timing (for profiling runs)")
    ap.add_argument("--no-skewed-run", action="store_true",
                    help="time only the skewed graph's sweeps, not its -c 8 -t 2 -a 0.25 run (for profiling runs)")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch tuning")
    ap.add_argument("--batch-graphs", type=int, default=4096, help="graphs in the batched-form run (0 = skip)")
    ap.add_argument("--batch-sample", type=int, default=256,
                    help="graphs of the batch also run one graph per launch (the comparison)")
    ap.add_argument("--skewed-n", type=int, default=1 << 22, help="BA vertices of the skewed-degree run (0 = skip)")
    ap.add_argument("--skewed-tuning", type=json.loads, default=None, metavar="JSON",
                    help="launch tuning of the skewed run, as the bench line's skewed.tuning object, instead of "
                         "the autotuner's (profiling passes then all run the same kernels)")
    ap.add_argument("--serve-requests", type=int, default=200, help="requests in the serving run (0 = skip)")
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--serve-load", type=float, default=0.9,
                    help="offered load as a fraction of the sequential solver's measured capacity")
    ap.add_argument("--probe-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB five times (counter calibration)")
    ap.add_argument("--probe-gather", action="store_true",
                    help="only time plain random 4-byte gathers curr[tail] over the skewed run's BA graph "
                         "(the floor of its local move's neighbour-community reads)")
    args = ap.parse_args()

    if args.probe_gather:
        print(json.dumps(gather_probe(args)), flush=True)
        return

    if args.probe_copy:
        import torch
        n = args.probe_copy * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"probe": "copy", "bytes_read_per_launch": 8 * n, "bytes_written_per_launch": 8 * n,
                          "launches": 5}), flush=True)
        return

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the path is single-device; multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import vite_amd

    ctx = vite_amd.Context(0)
    t0 = time.perf_counter()
    dg = ctx.gen_rgg(args.seed, args.n, args.extra)
    gen_s = time.perf_counter() - t0
    V, E = dg.nv, dg.ne
    log(f"graph: n={V} directed entries={E} generated in {gen_s:.2f}s")

    tuning = ctx.tuning
    default_tuning = tuning
    if not args.no_autotune:
        t0 = time.perf_counter()
        tuning = ctx.autotune(dg, prep=args.prep, reps=3)
        log(f"autotune: {tuning} in {time.perf_counter() - t0:.2f}s")
    st = dg.phase_begin(lower=-1.0, thresh=1e-3)
    for _ in range(args.prep):
        q, done = st.step()
        log(f"prep iteration: q={q:.9f}")
    for _ in range(args.warmup):
        st.sweep()
    ctx.set_profiling(True)
    ctx.reset_stats()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st.sweep()
    wall = time.perf_counter() - t0
    ms_per_step = 1e3 * wall / args.steps
    kavg, kcount, ktot = ctx.kernel_stats("local_move")
    ctx.set_profiling(False)
    state = st.state()
    per_iter = None
    if not args.no_iter:
        per_iter = iteration_timing(dg, st, V, E)
        log(f"whole iterations: {per_iter['ms_per_iteration']:.3f} ms each "
            f"({100 * per_iter['hbm_frac']:.1f}% of HBM peak in algorithmic bytes)")
    st.end(want_cvect=False)

    # the dominant kernel's own share: the rows of bucket 2 under the tuning in force
    deg = np.diff(dg.offsets())
    fused = degree_buckets(deg, tuning["bucket_max_degree"]) == G16_BUCKET
    Vf, Ef = int(np.count_nonzero(fused)), int(deg[fused].sum())
    alg_bytes = 24 * Ef + 56 * Vf
    achieved = alg_bytes / (kavg * 1e-3) / 1e9 if kavg > 0 else 0.0
    value = E / (ms_per_step * 1e-3)
    log(f"sweep: {ms_per_step:.3f} ms/step wall, kernel {kavg:.3f} ms avg over {kcount} launches "
        f"({Vf} of {V} rows, {Ef} of {E} entries), {achieved:.1f} GB/s algorithmic "
        f"({100 * achieved / HBM_PEAK_GBS:.1f}% of peak), {value / 1e9:.2f} G edges/s")
    tag = f"rgg_n{V}"
    traffic, traffic_src = measured_traffic("local_move", tag, g16_kernel(tuning))

    out = {
        "metric": METRIC, "value": value, "unit": "edges/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"RGG n={V} r=sqrt(2 ln n/(pi n)) +{args.extra}% extra edges U(0,1), "
                               f"local-move sweep after {args.prep} phase-0 iterations",
                   "n": V, "directed_entries": E, "seed": args.seed, "ids": "int64", "weights": "f64",
                   "threshold": 1e-3, "tuning": tuning},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": g16_kernel(tuning), "kernel_avg_ms": kavg, "algorithmic_bytes_per_launch": alg_bytes,
                     "bytes_formula": "24E + 56V over the kernel's rows (bucket 2: 16-lane groups)",
                     "sweep_achieved": (24 * E + 56 * V) / (ms_per_step * 1e-3) / 1e9,
                     "sweep_formula": "24E + 56V over all rows / wall time of the whole sweep (every launch)",
                     "kernel_rows": Vf, "kernel_entries": Ef,
                     "traffic_source": traffic_src,
                     "traffic_formula": f"FETCH_SIZE x {FETCH_SCALE} + WRITE_SIZE x {WRITE_SCALE} (KiB per dispatch)"},
        "per_iteration": per_iter,
    }

    if not args.no_ttq:
        # two whole runs on the same graph (the first also makes the context's allocations
        # for every level); the faster one is reported, both are listed.  They use the
        # library's default launch tuning: the autotuner times the level-0 sweep alone, and its
        # pick (a 2048-workgroup cap on this graph) makes the coarse levels slower.
        ctx.set_tuning(**default_tuning)
        runs = []
        for _ in range(2):
            t0 = time.perf_counter()
            r = dg.run(threshold_cycling=True)
            runs.append((time.perf_counter() - t0, r))
        ctx.set_tuning(**tuning)
        ttq, r = min(runs, key=lambda x: x[0])
        same = all(x.call_q == r.call_q and x.final_q == r.final_q for _, x in runs)
        out["time_to_final_q"] = {"seconds": ttq, "runs_seconds": [t for t, _ in runs], "runs_identical": same,
                                  "final_q": r.final_q, "levels": len(r.level_q),
                                  "calls": len(r.call_q), "total_iters": r.total_iters,
                                  "clustering_seconds": r.clustering_seconds, "flags": "-i",
                                  "tuning": default_tuning}
        ref = reference_run("rgg", args.seed, args.n, args.extra, ["-i"])
        if ref:
            out["time_to_final_q"]["reference"] = compare_reference(
                r, ref, "the reference's main() -i on the same graph, 1 MPI rank, run once in the build "
                        f"container ({ref['threads']} OpenMP threads), NOT on this host; the same-host "
                        "comparison is time_to_final_q_config1")
        log(f"time to final Q: {ttq:.2f}s q={r.final_q:.9f} levels={len(r.level_q)} iters={r.total_iters}")

    if args.skewed_n > 0:
        out["skewed"] = skewed_form(ctx, args)
    if args.batch_graphs > 0:
        out["batched"] = batched_form(ctx, args)
    if args.serve_requests > 0:
        out["serving"] = serving_form(args)

    if not args.no_cpu:
        host = host_cpus()
        if args.cpu_threads:
            host["threads_used"] = args.cpu_threads
        threads = host["threads_used"]
        log(f"cpu baseline: {threads} threads ({host}) ...")
        out["cpu_baseline"] = cpu_baseline(dg, state, threads, args.cpu_reps, state["target"], host)
        log(f"cpu baseline: {out['cpu_baseline']['value'] / 1e6:.1f} M edges/s, "
            f"mismatches {out['cpu_baseline']['target_mismatches_vs_gpu']}")
        if not args.no_ttq:
            out["time_to_final_q_config1"] = config1_time_to_q(ctx, args, host)
            log(f"config-1 time to final Q: {out['time_to_final_q_config1']}")
    dg.free()
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
