"""Benchmark of the pairwise Similarity path on one MI355X.

Workload (the headline configuration): a synthetic 20,000-gene x 500-sample
FP32 expression matrix from the module generator (10% NaN, seed 3), KINC
Similarity with GMM clustering (K = 1..5 by BIC), pre- and post-clustering
outlier removal, minimum 30 samples and Spearman correlation.  One step is
one KINC work block of 32,768 consecutive gene pairs; the steps sample the
199,990,000-pair space at every 97th block.  The expression matrix is
resident in HBM before timing starts and outputs stay on the device.

The JSON line reports pairs/s of the batched persistent form on `--workers`
engine handles taking blocks in turn (`value`, default 2), of the plain
one-launch-per-ksize form on one handle (`plain_value`) and on the same
handles (`plain_value_same_workers`), the FP32 VALU roofline
fraction of the dominant pair kernel, timed with HIP events on the stream it
runs on, and the CPU oracle timed on a bounded sample of the same pairs on
this host (which also re-checks bit-exact parity on that sample).

A serving leg follows: a Poisson stream of gene-set queries against a
resident 60,000 x 475 matrix (GMM + Pearson, |r| >= 0.5), with p50/p99
latency beside pairs/s, and the latency of one 100-gene query alone; then the
RMT threshold sweep on a 5,000-gene correlation matrix (and on a 700-gene one
beside the CPU oracle).

Usage: python bench.py [--gpus 1] [--steps 8] [--warmup 1] [--no-serve]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "gene pairs/sec GMM+Spearman 20k×500, batched vs plain; p50/p99 query latency"
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0
CLOCK_HZ = 2.4e9   # MI355X peak engine clock
WORKLOADS = {
    "H": "KINC similarity, GMM (K=1..5, BIC) + Spearman, pre/post outliers, minsamp 30; "
         "20000 genes x 500 samples FP32, module generator, 10% NaN",
    "B": "KINC similarity, GMM (K=1..5, BIC) + Spearman, pre/post outliers, minsamp 30; "
         "5000 genes x 500 samples FP32, module generator, 10% NaN",
    "A": "KINC similarity, no clustering, Pearson, minsamp 30; 1000 genes x 200 samples FP32, "
         "3-factor generator, 5% NaN",
    "C": "KINC similarity, K-means (K=1..5, BIC) + Pearson, minsamp 30; 20000 genes x 1000 "
         "samples FP32, module generator, per-gene NaN rate Beta(1,4)",
    "D": "KINC similarity, GMM (K=1..5, BIC) + Pearson, minsamp 30; 60000 genes x 475 samples "
         "FP32, module generator, 8% NaN (the serving matrix, as whole work blocks)",
    "E": "KINC similarity, GMM (K=1..5, BIC) + Spearman, minsamp 30; 2000 genes x 5000 samples "
         "FP32, module generator, 10% NaN, 20% of genes near-degenerate (precision stress)",
}
# HBM bytes per launch of the dominant pair kernel (k_pairs<7,5,4>, the
# 32,443-pair bin of block 97, one handle), from the FETCH_SIZE and
# WRITE_SIZE passes (scaled by a copy kernel of known size) recorded in:
TRAFFIC_SOURCE = "profiles/k_pairs_gmm_spearman_20000x500_counters.md"
TRAFFIC_BYTES_PER_LAUNCH = 9.98e9
# VALU wave64 instructions per second of the same dispatch (SQ_INSTS_VALU /
# its 40.68 ms) against the fastest VALU stream the issue probe measured on
# the same kind of box (v_add_f32, 4 waves per SIMD; profiles/valu_issue_calibration.md),
# and the lane utilisation SQ_THREAD_CYCLES_VALU / (64 SQ_ACTIVE_INST_VALU)
VALU_ISSUE_RATE = 8.15e11
VALU_ISSUE_CEILING = 1.17e12
VALU_LANE_UTILIZATION = 0.859
BLOCK = 32768


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pair_flops(N, iters, corr_method, labels, K, clus="gmm"):
    """Algorithmic FP32 flops of one pair from its per-fit iteration records.

    GMM fit with K components on N points, T E-steps and T_M M-steps:
    N*K*(18*T + 20*T_M) + 2*N*T (every add, sub, mul, div counts 1).
    Pearson adds 8 flops per in-cluster point; Spearman adds none (integer
    ranks).
    """
    total = 0
    for k in range(iters.shape[0]):
        it = int(iters[k])
        if it == 0:
            continue
        T = it & 0xFFF
        Kc = k + 1
        if clus == "kmeans":
            # per Lloyd iteration: K distances (2 sub, 2 mul, 2 add) per point
            # and the member sums (2 adds per point)
            total += N * T * (6 * Kc + 2)
            continue
        conv = (it >> 12) & 1
        TM = T - 1 if conv else T
        total += N * Kc * (18 * T + 20 * TM) + 2 * N * T
    if corr_method == "pearson":
        for k in range(int(K)):
            total += 8 * int(np.count_nonzero(labels == k))
    return total


def block_flops(K, labels, iters, corr_method, clus="gmm"):
    N = labels.shape[1] - np.count_nonzero((labels == -9) | (labels == -6), axis=1)
    f = np.zeros(K.shape[0], np.float64)
    for q in range(K.shape[0]):
        f[q] = pair_flops(int(N[q]), iters[q], corr_method, labels[q], K[q], clus)
    return f, N


def block_other_work(K, labels, iters, corr_method, minsamp, clus="gmm"):
    """Work the flop count leaves out: GMM exponentials N*K*(T + 2*T_M) and
    logarithms N*(T + 1) per fit, and the Spearman sorts' compare-exchanges,
    2 * (n2/2) * L(L+1)/2 per cluster of n >= minsamp points with
    n2 = next power of two of n (at least 2) and L = log2(n2)."""
    N = labels.shape[1] - np.count_nonzero((labels == -9) | (labels == -6), axis=1)
    exps = logs = cx = 0
    for q in range(K.shape[0]):
        if clus == "gmm":
            for k in range(iters.shape[1]):
                it = int(iters[q, k])
                if it == 0:
                    continue
                T = it & 0xFFF
                TM = T - 1 if (it >> 12) & 1 else T
                exps += int(N[q]) * (k + 1) * (T + 2 * TM)
                logs += int(N[q]) * (T + 1)
        if corr_method == "spearman":
            for k in range(int(K[q])):
                n = int(np.count_nonzero(labels[q] == k))
                if n < minsamp:
                    continue
                n2 = 2
                while n2 < n:
                    n2 *= 2
                L = n2.bit_length() - 1
                cx += 2 * (n2 // 2) * L * (L + 1) // 2
    return exps, logs, cx


def slots_for(n, S=500):
    """Register-slot class of a pair with n valid samples (the persistent
    form's bin; classes above the one S needs are not built)."""
    for p in (1, 2, 3, 4, 5, 6, 7, 8, 10, 12, 14, 16):
        if 64 * p >= n or 64 * p >= S:
            return p
    return 0


def run_form(native, x, flags, form, blocks, warmup, torch, workers=2, autotune=False):
    """Time the blocks on `workers` engine handles taking blocks in turn (the
    reference's several workers per device): each handle has its own stream,
    so a block's bins start while the previous block's last pairs finish.
    With autotune, the first handle times its launch configurations on the
    first 8192 pairs of the first block (kinc_sim_autotune) and every handle
    takes the fastest; untimed, before the warm-up."""
    import ctypes  # noqa: F401
    G, S = x.shape
    total_pairs = native.pair_count(G)
    p = native.make_params(form=form, **flags)
    engs = [native.Engine(x, p) for _ in range(workers)]
    tuning = engs[0].tuning()
    if autotune:
        start = blocks[0] * BLOCK
        # one work block run by all worker handles at once, as they share the
        # device in the timed region
        tuning = native.autotune_group(engs, start, min(BLOCK, total_pairs - start))
        log("  %s autotune: %s" % (form, tuning))
    MC = p.max_clusters
    nsteps = len(blocks)
    dev = torch.device("cuda:0")
    dK = torch.empty((nsteps, BLOCK), dtype=torch.int8, device=dev)
    dL = torch.empty((nsteps, BLOCK, S), dtype=torch.int8, device=dev)
    dC = torch.empty((nsteps, BLOCK, MC), dtype=torch.float32, device=dev)
    dI = torch.empty((nsteps, BLOCK, MC), dtype=torch.int16, device=dev)

    def run(i, b):
        start = b * BLOCK
        count = min(BLOCK, total_pairs - start)
        engs[i % workers].execute_device(start, count, dK[i].data_ptr(), dL[i].data_ptr(),
                                         dC[i].data_ptr(), dI[i].data_ptr())
        return count

    for i in range(max(warmup, workers)):
        run(i % nsteps, blocks[i % nsteps])
    for e in engs:
        e.synchronize()
    torch.cuda.synchronize()
    for e in engs:
        e.set_profiling(True)
        e.profile_read()
    t0 = time.perf_counter()
    pairs = 0
    for i, b in enumerate(blocks):
        pairs += run(i, b)
        log("  %s step %d/%d queued (block %d)" % (form, i + 1, nsteps, b))
    for e in engs:
        e.synchronize()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    prof = {}
    for e in engs:
        for P, (n, ms) in e.profile_read().items():
            n0, ms0 = prof.get(P, (0, 0.0))
            prof[P] = (n0 + n, ms0 + ms)
        e.set_profiling(False)
    out = dict(
        K=dK.cpu().numpy(), labels=dL.cpu().numpy(), corr=dC.cpu().numpy(),
        iters=dI.cpu().numpy().view(np.uint16))
    for e in engs:
        e.close()
    out["tuning"] = tuning
    return pairs, t1 - t0, prof, out


def census_leg(native, x, flags, blocks, out):
    """Phase census of the 7-slot, 5-cluster GMM instance on the timed blocks,
    after the timed region, on one handle: shader-clock cycles per phase summed
    over waves, exact counts of chain-sum instructions and of their active
    lanes, and a bit-exact check of the census build's results against the
    timed run's.  The census build reads the clock at every phase boundary,
    so its kernel time is reported beside the product build's on the same
    blocks."""
    import torch
    G, S = x.shape
    total = native.pair_count(G)
    p = native.make_params(form="persistent", **flags)
    e = native.Engine(x, p)
    e.set_tuning(4, 4)
    ctr = torch.zeros(native.CENSUS_COUNTERS, dtype=torch.int64, device="cuda:0")
    try:
        e.set_census(ctr.data_ptr())
    except native.KincError as err:
        e.close()
        return {"skipped": str(err)}
    e.set_profiling(True)
    e.profile_read()
    same = True
    for i, b in enumerate(blocks):
        start = b * BLOCK
        count = min(BLOCK, total - start)
        K, L, C, I = e.execute(start, count)
        same = same and np.array_equal(K, out["K"][i][:count]) and \
            np.array_equal(L, out["labels"][i][:count]) and \
            np.array_equal(I, out["iters"][i][:count]) and \
            np.array_equal(C.view(np.uint32), out["corr"][i][:count].view(np.uint32))
    prof_c = e.profile_read()
    e.set_census(None)
    for i, b in enumerate(blocks):
        start = b * BLOCK
        e.execute(start, min(BLOCK, total - start))
    prof_p = e.profile_read()
    e.set_profiling(False)
    e.close()
    c = native.census_dict(ctr.cpu().numpy())
    cyc = c["cycles"]
    wave = max(cyc["wave"], 1)
    named = sum(v for k, v in cyc.items() if k != "wave") + c["queue_cycles"]
    pairs = max(c["pairs"], 1)
    res = {
        "kernel": "k_pairs_census<P=7,KMAX=5> (instrumented build of k_pairs<7,5,4>)",
        "results_identical_to_timed_run": bool(same),
        "kernel_ms": {"census": prof_c.get(7, (0, 0.0))[1], "product": prof_p.get(7, (0, 0.0))[1]},
        "pairs": c["pairs"],
        "cycle_share": dict({k: v / wave for k, v in cyc.items() if k != "wave"},
                            queue=c["queue_cycles"] / wave, unattributed=(wave - named) / wave),
        "chain_sum_share_of_cycles": (cyc["a_chain"] + cyc["b_chain"]) / wave,
        "chain_lane_utilization": c["chain_lanes"] / max(64 * c["chain_adds"], 1),
        "per_pair": {k: c[k] / pairs for k in ("fits", "iters", "estep_slots", "a_slots",
                                               "b_slots", "chain_adds", "chain_calls", "exps",
                                               "logs")},
        "counters": c,
    }
    return res


def cpu_baseline(oracle, native, x, flags, blocks, seconds, gpu_out):
    """Oracle (single thread) on a prefix of every timed block, sized so the
    whole sample takes about `seconds` of CPU time; each prefix starts at its
    block's first pair, so block-dependent semantics (Pearson's shared X
    workspace) match the device run.  Returns pairs/s and the bit-exact check
    of the sampled pairs against the device outputs."""
    from parity_util import count_mismatching_pairs as compare_block
    p = native.make_params(**flags)
    done = 0
    t_used = 0.0
    mismatches = 0
    errs = []
    wrapped = 0
    per_block = 16
    for i, b in enumerate(blocks):
        if i == 1:
            # size the remaining prefixes from the rate seen on the first one
            rate = done / max(t_used, 1e-6)
            per_block = int(max(8, min(BLOCK, rate * (seconds - t_used) / max(1, len(blocks) - 1))))
        start = b * BLOCK
        t0 = time.process_time()
        want = oracle.execute_block(p, x, start, per_block)
        t_used += time.process_time() - t0
        done += per_block
        got = tuple(a[i][:per_block] for a in
                    (gpu_out["K"], gpu_out["labels"], gpu_out["corr"], gpu_out["iters"]))
        mismatches += compare_block(got, want)
        # the device's correlations against the same clusters' correlations
        # in FP64 (the oracle's double-precision Pearson / exact-integer
        # Spearman)
        f64 = oracle.corr_fp64_block(p, x, start, per_block, got[0], got[1])
        ok = np.isfinite(f64) & np.isfinite(got[2])
        if flags.get("corrmethod") == "spearman":
            # clusters of 1,291 or more points: the reference's int32
            # n*(n*n - 1) wraps (pairwise_spearman.cpp:73-82), reproduced
            # here, so the FP64 value is not the same quantity
            sizes = np.stack([(got[1] == k).sum(axis=1) for k in range(f64.shape[1])], axis=1)
            wrap = ok & (sizes >= 1291)
            wrapped += int(wrap.sum())
            ok &= ~wrap
        errs.append(np.abs(got[2][ok].astype(np.float64) - f64[ok]))
    e = np.concatenate(errs) if errs else np.zeros(0)
    fp64 = {"max_abs": float(e.max()) if e.size else None,
            "mean_abs": float(e.mean()) if e.size else None, "clusters": int(e.size),
            "excluded_int32_wrap": wrapped}
    return done / t_used, done, t_used, done, mismatches, fp64


def cpu_model():
    """The host CPU's model name (lscpu's "Model name")."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def host_threads():
    """Host threads the all-core CPU leg may use: the affinity set, at most 16
    (a one-GPU share of the host)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline_threads(oracle, native, x, flags, blocks, seconds, gpu_out, rate1, threads):
    """The oracle on every host thread at once (the reference parallelises by
    work blocks), sized from the single-thread rate so the leg takes about
    `seconds` of wall time.  Each thread takes a run of pairs from the start
    of a timed block; without Pearson (whose stale workspace ties a pair to
    its block) a block may give several consecutive runs, one per thread, so
    every host thread has work.  The device outputs check every pair bit for
    bit.  Returns pairs/s (wall), pairs, wall seconds, mismatches, threads."""
    from concurrent.futures import ThreadPoolExecutor
    from parity_util import count_mismatching_pairs as compare_block
    p = native.make_params(**flags)
    per_run = int(max(8, min(BLOCK, rate1 * seconds)))
    runs_per_block = 1
    if flags.get("corrmethod", "pearson") != "pearson":
        runs_per_block = max(1, -(-threads // len(blocks)))
        per_run = min(per_run, BLOCK // runs_per_block)
    tasks = [(i, b, r * per_run) for i, b in enumerate(blocks) for r in range(runs_per_block)]
    tasks = tasks[:threads]

    def work(t):
        i, b, off = t
        return i, off, oracle.execute_block(p, x, b * BLOCK + off, per_run)

    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=len(tasks)) as ex:
        results = list(ex.map(work, tasks))
    wall = time.perf_counter() - t0
    mismatches = 0
    for i, off, want in results:
        got = tuple(a[i][off:off + per_run] for a in
                    (gpu_out["K"], gpu_out["labels"], gpu_out["corr"], gpu_out["iters"]))
        mismatches += compare_block(got, want)
    done = per_run * len(tasks)
    return done / wall, done, wall, mismatches, len(tasks)


def serve_leg(native, gemgen, n_queries, qps, max_genes, big_query, seed=11):
    """Config D's serving load: a Poisson query stream against the resident
    60,000 x 475 matrix (GMM + Pearson, |r| >= 0.5), query sizes log-uniform
    on [1, max_genes] genes, all partners of each query gene; latency is
    submit-to-answer as measured by the server, and start wait the part of it
    before the query's first work block was launched.  Then a 1-gene query
    submitted while a big_query-gene query runs shows how long a small query
    waits behind a large one (at most the two one-block batches in flight).
    Returns the `serving` object of the JSON line."""
    import threading
    from kinc_amd import serving
    cfg = gemgen.CONFIGS["D"]
    t = time.time()
    x = gemgen.make("D")
    G = x.shape[0]
    log("serving: generated %dx%d GEM in %.1fs" % (G, x.shape[1], time.time() - t))
    p = native.make_params(**cfg["params"])
    quantum = 32768
    srv = serving.Server(x, p, hbm_budget=16 << 30, batch_pairs=quantum)
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(0.0, np.log(max_genes), n_queries)).astype(int).clip(1, max_genes)
    queries = [rng.choice(G, size=int(k), replace=False) for k in sizes]
    gaps = rng.exponential(1.0 / qps, n_queries)
    ids = []
    lock = threading.Lock()

    def producer():
        for qgenes, gap in zip(queries, gaps):
            time.sleep(gap)
            qid = srv.submit(qgenes, 0.5)
            with lock:
                ids.append(qid)

    p0, b0, busy0 = srv.stats()
    t0 = time.perf_counter()
    th = threading.Thread(target=producer)
    th.start()
    lat, waits, npairs, nhits, per_query = [], [], 0, 0, []
    done = 0
    while done < n_queries:
        with lock:
            pending = ids[done:]
        if not pending:
            time.sleep(0.01)
            continue
        qid = pending[0]
        assert srv.wait(qid, 1200)
        w, r = srv.timing(qid)
        hits, l, n = srv.result(qid)
        lat.append(l)
        waits.append(w)
        npairs += n
        nhits += len(hits)
        per_query.append({"genes": int(sizes[done]), "pairs": int(n), "latency_s": l,
                          "start_wait_s": w})
        done += 1
        log("  query %d/%d: %d genes, %d pairs, %d hits, %.3f s (start wait %.3f s)"
            % (done, n_queries, sizes[done - 1], n, len(hits), l, w))
    th.join()
    wall = time.perf_counter() - t0
    p1, b1, busy1 = srv.stats()
    out = {
        "workload": "60000 genes x 475 samples FP32 (module generator, 8% NaN), GMM K=1..5 BIC + "
                    "Pearson, |r| >= 0.5, all partners of each query gene",
        "queries": n_queries, "qps": qps, "query_genes": "log-uniform on [1, %d]" % max_genes,
        "quantum_pairs": quantum,
        "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
        "max_s": float(np.max(lat)),
        "start_wait_p50_s": float(np.percentile(waits, 50)),
        "start_wait_p99_s": float(np.percentile(waits, 99)),
        "pairs_per_s": (p1 - p0) / wall, "device_busy_frac": (busy1 - busy0) / wall,
        "pairs": int(npairs), "hits": int(nhits), "batches": int(b1 - b0),
        "per_query": per_query,
    }
    if big_query:
        qbig = srv.submit(rng.choice(G, size=big_query, replace=False), 0.5)
        time.sleep(1.0)
        qone = srv.submit(rng.choice(G, size=1, replace=False), 0.5)
        assert srv.wait(qone, 1200)
        w1, r1 = srv.timing(qone)
        _, l1, n1 = srv.result(qone)
        assert srv.wait(qbig, 1200)
        wb, rb = srv.timing(qbig)
        hb, lb, nb = srv.result(qbig)
        log("  %d-gene query: %d pairs, %.2f s; 1-gene query submitted 1 s into it: %d pairs, "
            "%.3f s (start wait %.3f s)" % (big_query, nb, lb, n1, l1, w1))
        out["single_query"] = {"genes": big_query, "pairs": int(nb), "latency_s": lb}
        out["small_behind_large"] = {"genes": 1, "pairs": int(n1), "latency_s": l1,
                                     "start_wait_s": w1, "run_s": r1,
                                     "behind": "%d-gene query submitted 1 s earlier" % big_query}
    srv.close()
    return out


def graph_leg(native, gemgen, torch, reps=20):
    """Config A (1,000 x 200, Pearson, no clustering: 16 work blocks of a
    few launches each, so launch gaps matter most) on one handle, eager
    (every block's kernels enqueued by the host) against the same blocks
    recorded once into a HIP graph and replayed; best of `reps` passes over
    all blocks each, outputs checked equal after every replay."""
    x = gemgen.make("A")
    G, S = x.shape
    cfg = gemgen.CONFIGS["A"]
    p = native.make_params(**cfg["params"])
    total = native.pair_count(G)
    starts = list(range(0, total, BLOCK))
    n = len(starts)
    dev = torch.device("cuda:0")
    MC = p.max_clusters

    def outputs():
        return (torch.empty((n, BLOCK), dtype=torch.int8, device=dev),
                torch.empty((n, BLOCK, S), dtype=torch.int8, device=dev),
                torch.empty((n, BLOCK, MC), dtype=torch.float32, device=dev),
                torch.empty((n, BLOCK, MC), dtype=torch.int16, device=dev))

    eng = native.Engine(x, p)
    try:
        out = outputs()

        def eager_pass():
            for i, st in enumerate(starts):
                eng.execute_device(st, min(BLOCK, total - st), *(t[i].data_ptr() for t in out))

        # the same fill before the eager pass and before every replay: rows
        # past the last block's pairs are written by neither
        for t in out:
            t.fill_(-3)
        eager_pass()
        eng.synchronize()
        want = [t.clone() for t in out]
        best_eager = 1e9
        for _ in range(reps):
            t0 = time.perf_counter()
            eager_pass()
            eng.synchronize()
            best_eager = min(best_eager, time.perf_counter() - t0)
        eng.graph_begin()
        eager_pass()
        g = eng.graph_end()
        best_graph = 1e9
        same = True
        for _ in range(reps):
            for t in out:
                t.fill_(-3)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            g.launch(None)
            torch.cuda.synchronize()
            best_graph = min(best_graph, time.perf_counter() - t0)
            # bytes, so that NaN correlations compare equal to themselves
            same = same and all(torch.equal(a.view(torch.uint8), b.view(torch.uint8))
                                for a, b in zip(out, want))
        g.close()
        host = [tuple(a[i, :min(BLOCK, total - st)].cpu().numpy() for a in want)
                for i, st in enumerate(starts)]
    finally:
        eng.close()
    res = {"workload": WORKLOADS["A"], "blocks": n, "pairs": int(total),
           "eager_pairs_per_s": total / best_eager, "graph_pairs_per_s": total / best_graph,
           "eager_ms": best_eager * 1e3, "graph_ms": best_graph * 1e3,
           "replays_identical": bool(same), "handles": 1}
    log("config A on one handle: eager %.3f ms, graph %.3f ms per pass over %d blocks; "
        "replays identical: %s" % (best_eager * 1e3, best_graph * 1e3, n, same))
    res["_device_outputs"] = (x, p, starts, host)
    return res


def cpu_baseline_config_a(oracle, graph_obj):
    """Config A timed in full on the CPU oracle (one thread, then one thread
    per work block), every pair checked against the device outputs of the
    graph leg."""
    from concurrent.futures import ThreadPoolExecutor
    from parity_util import count_mismatching_pairs
    x, p, starts, host = graph_obj["_device_outputs"]
    total = graph_obj["pairs"]
    t0 = time.perf_counter()
    cpu = [oracle.execute_block(p, x, st, min(BLOCK, total - st)) for st in starts]
    t1 = time.perf_counter() - t0
    threads = host_threads()
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as ex:
        cpu_t = list(ex.map(lambda st: oracle.execute_block(p, x, st, min(BLOCK, total - st)),
                            starts))
    tn = time.perf_counter() - t0
    bad = sum(count_mismatching_pairs((h[0], h[1], h[2], h[3].view(np.uint16)), c)
              for h, c in zip(host, cpu))
    bad += sum(count_mismatching_pairs(a, b) for a, b in zip(cpu, cpu_t))
    log("config A in full on the CPU oracle: %.0f pairs/s on 1 thread, %.0f on %d; %d "
        "mismatching pairs" % (total / t1, total / tn, threads, bad))
    return {"workload": WORKLOADS["A"], "pairs": int(total), "kind": "port",
            "single_thread_pairs_per_s": total / t1, "threads": threads,
            "threads_pairs_per_s": total / tn, "mismatching_pairs": int(bad),
            "cpu_model": cpu_model(), "host_cores": os.cpu_count()}


def rmt_leg(gemgen, with_cpu):
    """The RMT threshold sweep (include/kinc_rmt.h) on a 5,000-gene matrix
    of 1,000 independent pairs and a 3,000-gene correlated block, and on the
    700-gene version of it beside the CPU oracle (Jacobi eigensolver)."""
    from kinc_amd import rmt
    out = {}
    m = gemgen.rmt_pairs_and_block(1000, 3000, seed=2)
    t0 = time.perf_counter()
    h = rmt.Rmt.from_dense(m)
    t1 = time.perf_counter()
    rc, thr, rows = h.run()
    t2 = time.perf_counter()
    if rc != 0:
        log("rmt: %s" % h.last_error())
    h.close()
    out.update({
        "workload": "RMT threshold sweep from 0.99 by 0.001 on 5000 genes (1000 independent "
                    "pairs r~U(0.55,0.99) + a 3000-gene block r=0.5405+-2e-4), 1 cluster",
        "records": int(np.count_nonzero(m) // 2), "thresholds": len(rows),
        "final_threshold": thr, "largest_pruned": max(r[1] for r in rows),
        "upload_s": t1 - t0, "sweep_s": t2 - t1,
        "ms_per_threshold": (t2 - t1) / max(1, len(rows)) * 1e3})
    log("rmt: %d thresholds in %.2fs (final %.4f, largest %d)"
        % (len(rows), t2 - t1, thr, out["largest_pruned"]))
    small = gemgen.rmt_pairs_and_block(100, 500, seed=1)
    h = rmt.Rmt.from_dense(small)
    t0 = time.perf_counter()
    rc, thr_s, rows_s = h.run()
    gpu_s = time.perf_counter() - t0
    h.close()
    out["small"] = {"genes": 700, "thresholds": len(rows_s), "sweep_s": gpu_s,
                    "final_threshold": thr_s}
    if with_cpu:
        import oracle
        t0 = time.process_time()
        orc, othr, orows = oracle.rmt_run(small)
        cpu_s = time.process_time() - t0
        out["small"]["cpu_oracle"] = {"sweep_s": cpu_s, "cores": 1, "kind": "port",
                                      "final_threshold": othr,
                                      "same_threshold": othr == thr_s}
        log("rmt small: gpu %.2fs, cpu oracle %.2fs, thresholds %s / %s"
            % (gpu_s, cpu_s, thr_s, othr))
    return out


# VALU instructions per operation of the probe modes that run compiled code:
# one exponential, as the pair kernels compile it (kinc_expf_le0: one clamp,
# multiply, v_rndne, two reduction FMAs, six polynomial FMAs, v_cvt, v_ldexp;
# kinc_expf: a second clamp), counted in the probe loop's gfx950 assembly,
# which holds exactly 128 of them and no other VALU instruction
PROBE_VALU_PER_OP = {"exp_le0": 13, "exp": 14}


def valu_calibration():
    """Issue rate of VALU instruction streams with 16 one-wave workgroups per
    CU (4 waves per SIMD, as the pair kernels run): wave64 instructions per
    second and cycles per instruction per SIMD at the 2.4 GHz peak clock, for
    full, partial and dependent streams of each probed opcode; then one wave
    per SIMD, which shows the dependent-instruction latency."""
    import torch
    from kinc_amd import _native as native
    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    wgs = cus * 16
    lds = 160 * 1024 // 16
    iters = 20000
    rows = []
    for mode, (chains, lanes, op) in enumerate(native.VALU_PROBE_MODES):
        ms = native.valu_probe(mode, iters, wgs, lds)
        name = native.VALU_PROBE_OPS[op]
        insts = 128.0 * iters * wgs * PROBE_VALU_PER_OP.get(name, 1)
        rate = insts / (ms * 1e-3)
        cpi = (cus * 4) * CLOCK_HZ / rate
        rows.append({"mode": mode, "op": name, "chains": chains, "lanes": lanes, "ms": ms,
                     "wave_insts": insts, "insts_per_s": rate,
                     "cycles_per_inst_per_simd": cpi})
        log("valu probe mode %d (%s, %d chains, %d lanes): %.3f ms, %.3g wave insts/s, "
            "%.2f cycles per instruction per SIMD" % (mode, name, chains, lanes, ms, rate, cpi))
    # one wave per SIMD (4 per CU): the cycles between one wave's instructions,
    # dependent (one chain) and independent (eight chains)
    single = []
    for mode in (3, 0, 15, 7):
        chains, lanes, op = native.VALU_PROBE_MODES[mode]
        ms = native.valu_probe(mode, 2000, cus * 4, 0)
        cyc = ms * 1e-3 * CLOCK_HZ / (128.0 * 2000)
        single.append({"mode": mode, "op": native.VALU_PROBE_OPS[op], "chains": chains,
                       "lanes": lanes, "ms": ms, "cycles_per_inst_per_wave": cyc})
        log("one wave per SIMD, mode %d (%s, %d chains): %.2f cycles per instruction"
            % (mode, native.VALU_PROBE_OPS[op], chains, cyc))
    print(json.dumps({"valu_calibration": rows, "one_wave_per_simd": single, "cus": cus,
                      "workgroups": wgs, "iters": iters, "clock_hz": CLOCK_HZ}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-cpu-threads", action="store_true",
                    help="time the CPU oracle on one thread only")
    ap.add_argument("--config", default="H", choices=["A", "B", "H", "C", "D", "E"],
                    help="workload (H is the headline; the others are reported for comparison)")
    ap.add_argument("--workers", type=int, default=2,
                    help="engine handles taking blocks in turn on one GPU (each its own stream)")
    ap.add_argument("--set", action="append", default=[], metavar="FLAG=VALUE",
                    help="override one analytic flag of the workload (e.g. preout=false, minsamp=20)")
    ap.add_argument("--samples", type=int, default=None,
                    help="override the workload's sample count (same generator and seed)")
    ap.add_argument("--clusmethod", choices=["none", "gmm", "kmeans"], default=None,
                    help="override the workload's clustering method (config C runs both)")
    ap.add_argument("--pearson-fp64", action="store_true",
                    help="Pearson sums and quotient in double (the pearson_fp64 option)")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the handles' default launch configuration (no kinc_sim_autotune)")
    ap.add_argument("--no-rmt", action="store_true", help="skip the RMT threshold-sweep leg")
    ap.add_argument("--no-census", action="store_true",
                    help="skip the phase census of the 7-slot GMM kernel")
    ap.add_argument("--no-graph", action="store_true",
                    help="skip the config A eager-versus-HIP-graph leg")
    ap.add_argument("--serve-queries", type=int, default=8)
    ap.add_argument("--serve-qps", type=float, default=0.1)
    ap.add_argument("--serve-max-genes", type=int, default=500)
    ap.add_argument("--serve-big-query", type=int, default=100)
    ap.add_argument("--valu-calibration", action="store_true",
                    help="only run the VALU issue-rate probe (v_add_f32 streams, 4 waves/SIMD)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (HBM counter calibration)")
    args = ap.parse_args()

    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": n * 4, "launches": 3}))
        return

    if args.valu_calibration:
        valu_calibration()
        return

    if args.gpus != 1:
        print("bench.py: measures one GPU only; the pair space would shard "
              "across GPUs by contiguous work-block ranges (see the sharding notes); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from kinc_amd import _native as native, gemgen
    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)
    torch.cuda.set_device(0)
    native.lib()

    cfg = gemgen.CONFIGS[args.config]
    flags = dict(cfg["params"])
    if args.pearson_fp64:
        flags["pearson_fp64"] = True
    if args.clusmethod:
        flags["clusmethod"] = args.clusmethod
    for kv in args.set:
        key, _, val = kv.partition("=")
        if val.lower() in ("true", "false"):
            flags[key] = val.lower() == "true"
        else:
            try:
                flags[key] = int(val)
            except ValueError:
                try:
                    flags[key] = float(val)
                except ValueError:
                    flags[key] = val
    t = time.time()
    x = gemgen.make(args.config, S=args.samples)
    G, S = x.shape
    log("generated %dx%d GEM in %.1fs" % (G, S, time.time() - t))
    total_pairs = native.pair_count(G)
    nblocks = (total_pairs + BLOCK - 1) // BLOCK
    blocks = sorted({(97 * (i + 1)) % nblocks for i in range(args.steps)})
    if len(blocks) < args.steps:
        blocks = list(range(min(args.steps, nblocks)))

    pairs, secs, prof, out = run_form(native, x, flags, "persistent", blocks, args.warmup, torch,
                                      args.workers, autotune=not args.no_autotune)
    value = pairs / secs
    log("persistent: %d pairs in %.3fs -> %.0f pairs/s; kernels %s" % (pairs, secs, value, prof))

    plain_value = None
    plain_value_workers = None
    if not args.no_plain:
        # the plain form as the reference runs its blocks: one launch per
        # ksize sub-batch, one block after another on one handle
        pp, ps, pprof, pout = run_form(native, x, flags, "plain", blocks, args.warmup, torch, 1)
        plain_value = pp / ps
        log("plain: %d pairs in %.3fs -> %.0f pairs/s" % (pp, ps, plain_value))
        if args.workers > 1:
            # and with the same worker handles as the persistent form
            pw, pws, _, _ = run_form(native, x, flags, "plain", blocks, args.warmup, torch,
                                     args.workers)
            plain_value_workers = pw / pws
            log("plain, %d workers: %d pairs in %.3fs -> %.0f pairs/s"
                % (args.workers, pw, pws, plain_value_workers))
        same = all(np.array_equal(out[k], pout[k]) for k in ("K", "labels", "iters")) and \
            np.array_equal(np.nan_to_num(out["corr"], nan=9.0).view(np.uint32),
                           np.nan_to_num(pout["corr"], nan=9.0).view(np.uint32))
        log("plain == persistent (bit-exact):", same)

    # algorithmic flops per step and per pair class
    flops_total = 0.0
    flops_by_P = {}
    for i in range(len(blocks)):
        f, N = block_flops(out["K"][i], out["labels"][i], out["iters"][i], flags["corrmethod"],
                           flags.get("clusmethod", "none"))
        flops_total += f.sum()
        Ps = np.array([slots_for(int(n), S) for n in N])
        for P in np.unique(Ps):
            flops_by_P[int(P)] = flops_by_P.get(int(P), 0.0) + float(f[Ps == P].sum())
    # the dominant kernel is the bin instance that carries the most work; bins
    # run concurrently, so each one's duration spans the others' too
    clustered = flags.get("clusmethod", "none") != "none"
    if clustered:
        dom = max((P for P in flops_by_P if P in prof), key=lambda P: flops_by_P[P])
        launches, dom_ms = prof[dom]
        achieved = flops_by_P.get(dom, 0.0) / launches / (dom_ms / launches * 1e-3) / 1e12
        roof = {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS,
                "flops_formula": "GMM: N*K*(18T + 20T_M) + 2NT per fit; K-means: N*T*(6K + 2); "
                                 "Pearson: 8 per in-cluster point"}
    else:
        # without clustering one launch strides over the block: HBM-bound
        # (algorithmic bytes per pair 8*S + 1 + 4*maxclus)
        dom = max(prof, key=lambda P: prof[P][1])
        launches, dom_ms = prof[dom]
        per_pair = 8 * S + 1 + 4 * flags.get("maxclus", 5)
        achieved = per_pair * pairs / (dom_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS,
                "bytes_formula": "8*S + 1 + S*[K>1] + 4*maxclus per pair"}
    whole_tflops = flops_total / secs / 1e12
    other = [0, 0, 0]
    for i in range(len(blocks)):
        e, l_, c = block_other_work(out["K"][i], out["labels"][i], out["iters"][i],
                                    flags["corrmethod"], flags.get("minsamp", 30),
                                    flags.get("clusmethod", "none"))
        other[0] += e
        other[1] += l_
        other[2] += c
    other_work = {
        "gmm_exp_per_s": other[0] / secs, "gmm_log_per_s": other[1] / secs,
        "spearman_compare_exchanges_per_s": other[2] / secs,
        "per_pair": {"gmm_exp": other[0] / pairs, "gmm_log": other[1] / pairs,
                     "spearman_compare_exchanges": other[2] / pairs},
        "note": "software exp/log (15 and about 30 VALU instructions), counted apart from the flops",
    }
    # the counter profile covers the headline workload as generated
    profiled = (args.config == "H" and not args.samples and not args.clusmethod
                and not args.pearson_fp64 and not args.set)

    census_obj = None
    if clustered and flags.get("clusmethod") == "gmm" and not args.no_census:
        census_obj = census_leg(native, x, flags, blocks, out)
        log("census: %s" % json.dumps({k: v for k, v in census_obj.items() if k != "counters"}))

    graph_obj = None
    if not args.no_graph:
        graph_obj = graph_leg(native, gemgen, torch)

    cpu = None
    fp64_err = None
    if not args.no_cpu:
        import oracle
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        rate, n, used, checked, bad, fp64_err = cpu_baseline(oracle, native, x, flags, blocks,
                                                             args.cpu_seconds, out)
        log("correlations vs FP64 on the sampled pairs: max %.3g, mean %.3g over %d clusters"
            " (%d clusters with int32-wrapped Spearman sums excluded)"
            % (fp64_err["max_abs"] or 0, fp64_err["mean_abs"] or 0, fp64_err["clusters"],
               fp64_err["excluded_int32_wrap"]))
        log("cpu oracle: %d pairs in %.1fs cpu -> %.2f pairs/s; parity %d/%d mismatching"
            % (n, used, rate, bad, checked))
        cpu = {"value": rate, "unit": "pairs/s", "cores": 1, "kind": "port",
               "cpu_model": cpu_model(), "host_cores": os.cpu_count(),
               "sample": "%d pairs: a prefix of each of the %d timed work blocks, single-threaded "
                         "C oracle (-O2, no FP contraction); GPU output bit-exact on %d/%d"
                         % (n, len(blocks), checked - bad, checked)}
        if not args.no_cpu_threads:
            threads = host_threads()
            if threads > 1:
                mrate, mn, mwall, mbad, used_t = cpu_baseline_threads(
                    oracle, native, x, flags, blocks, args.cpu_seconds * 2 / 3, out, rate, threads)
                log("cpu oracle, %d threads: %d pairs in %.1fs wall -> %.2f pairs/s; parity %d/%d "
                    "mismatching" % (used_t, mn, mwall, mrate, mbad, mn))
                cpu = {"value": mrate, "unit": "pairs/s", "cores": used_t, "kind": "port",
                       "cpu_model": cpu_model(), "host_cores": os.cpu_count(),
                       "sample": "%d pairs: one run of %d pairs of a timed work block per thread, "
                                 "C oracle (-O2, no FP contraction) on %d host threads; GPU output "
                                 "bit-exact on %d/%d" % (mn, mn // used_t, used_t, mn - mbad, mn),
                       "single_thread": {"value": rate, "cores": 1, "sample": cpu["sample"]}}
        if graph_obj is not None:
            cpu["config_a_full"] = cpu_baseline_config_a(oracle, graph_obj)
    if graph_obj is not None:
        graph_obj.pop("_device_outputs", None)

    rmt_obj = None
    if not args.no_rmt:
        rmt_obj = rmt_leg(gemgen, not args.no_cpu)

    serving_obj = None
    if not args.no_serve:
        serving_obj = serve_leg(native, gemgen, args.serve_queries, args.serve_qps,
                                args.serve_max_genes, args.serve_big_query)

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": len(blocks),
        "warmup": args.warmup,
        "ms_per_step": secs / len(blocks) * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": WORKLOADS[args.config] + (
                "; clustering overridden: %s" % args.clusmethod if args.clusmethod else "") + (
                "; samples overridden: %d" % args.samples if args.samples else "") + (
                "; flags overridden: %s" % ", ".join(args.set) if args.set else ""),
            "pairs_per_step": BLOCK,
            "step": "one 32768-pair work block",
            "form": "persistent",
            "parallelism": "1 GPU",
            "workers": args.workers,
            "pearson_fp64": bool(args.pearson_fp64),
            "tuning": out.get("tuning"),
        },
        "plain_value": plain_value,
        "batched_vs_plain": (value / plain_value) if plain_value else None,
        "plain_value_same_workers": plain_value_workers,
        "batched_vs_plain_same_workers": (value / plain_value_workers) if plain_value_workers else None,
        "roofline": dict(roof, **{
            "kernel": ("k_pearson_lanes (one pair per lane, persistent form)" if dom == -1
                       else "k_pairs<P=%d,KMAX=5> (persistent form)" % dom),
            "traffic": TRAFFIC_BYTES_PER_LAUNCH if profiled else None,
            "traffic_source": TRAFFIC_SOURCE if profiled else None,
            "valu_issue": ({"insts_per_s": VALU_ISSUE_RATE, "ceiling": VALU_ISSUE_CEILING,
                            "frac": VALU_ISSUE_RATE / VALU_ISSUE_CEILING,
                            "ceiling_is": "fastest measured VALU stream (v_add_f32, 4 waves "
                                          "per SIMD)",
                            "lane_utilization": VALU_LANE_UTILIZATION,
                            "source": TRAFFIC_SOURCE} if profiled else None),
            "whole_job_tflops": whole_tflops,
        }),
        "other_work": other_work,
        "cpu_baseline": cpu,
        "corr_error_vs_fp64": fp64_err,
        "serving": serving_obj,
        "graph": graph_obj,
        "census": census_obj,
        "rmt": rmt_obj,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
