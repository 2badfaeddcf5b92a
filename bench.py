"""Benchmark of the MI355X database search (one GPU).

Default workload (BASELINE configuration "SW 8-bit with the overflow cascade"):
200,000 synthetic protein sequences (lognormal lengths, median 300, sigma 0.6,
cap 35,000; 5 % planted query fragments at 60-95 % identity), queries of 50,
500 and 5,000 residues, BLOSUM62, gaps open 10 / extend 1, top 50 hits, 8-bit
start with the 16/64-bit cascade, scores only.  One step = the three queries
searched through sw_align with the DB resident on the GPU.

    python bench.py [--steps K] [--warmup W] [--workload all|c1|c2|c3|c4|c5|dbload] [--no-cpu-baseline]

Prints one JSON line.  value = algorithmic cells (sum over queries and DB
entries of |q| * |d|) per second of the whole step, in GCUPS.  The default
(all) adds, beside that headline, the NW DNA workload (nw_dna), the batched
workload (batched: batched vs one query per launch) and the served mix
(served: p50/p99 and GCUPS at a sustainable rate and under overload, one
request per launch vs continuous batching), each at its full size.

c4 (batched): 256 queries (log-uniform 30-2000) against 500,000 sequences,
sw_align(q, 20, BIT_WIDTH_8, COMPUTE_SCORE) each, run one query per call and
then as one batch (ssa_amd_batch_align); value = batched GCUPS.
c5 (served): 1,000,000 sequences resident; Poisson arrivals of an 80/20 SW/NW
mix with CIGARs (top 32, 8-bit start), query lengths 70 % U(50,400) + 30 %
log-U(400,5000); the same arrival trace is served one request per launch and
with continuous batching; p50/p99 latency per rate.
"""
import argparse
import ctypes as C
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_VALU_TOPS = 256 * 64 * 2 * 2.4e9 / 1e12  # packed int16 element-ops/s: 256 CU x 64 lanes x 2 x 2.4 GHz
OPS_PER_CELL = {"sw": 9, "nw": 8}             # algorithmic int ops per cell of the reference recurrence
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "score_kernel_hbm_traffic.txt")


def measured_traffic(workload):
    """HBM bytes per step of the score kernels, as measured by the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE runs of this workload (None if absent)."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    with open(TRAFFIC_FILE) as f:
        for line in f:
            parts = line.split()
            if len(parts) == 3 and parts[0] == "bytes_per_step" and parts[1] == workload:
                return int(parts[2])
    return None


def workload(name):
    from libssa_amd import synthetic as S
    if name == "c2":
        qs, db, _ = S.config_c2()
        return dict(kind="sw", queries=qs, db=db, k=50, bw=8, symtype=1, strands=1, matrix="blosum62",
                    gaps=(-10, -1), desc="sw 8-bit cascade, 200k lognormal protein DB, q=50/500/5000, k=50")
    if name == "c1":
        qs, db, _ = S.config_c1()
        return dict(kind="sw", queries=qs, db=db, k=10, bw=16, symtype=1, strands=1, matrix="blosum62",
                    gaps=(-10, -1), desc="sw 16-bit, 2k uniform protein DB, q=120, k=10")
    if name == "c3":
        qs, db, _ = S.config_c3()
        return dict(kind="nw", queries=qs, db=db, k=100, bw=16, symtype=0, strands=3, const=(2, -4),
                    gaps=(-20, -2), align=True,
                    desc="nw 16-bit DNA both strands with CIGARs, 1M DB, q=250, k=100")
    raise SystemExit("unknown workload " + name)


def host_cpus():
    """The CPUs this process may use (its affinity set, at most 16: the host's
    CPU share per GPU), the physical cores among them, and the CPU model."""
    allowed = sorted(os.sched_getaffinity(0))
    logical = allowed[:16]
    model, core_of, cur = "unknown", {}, {}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f.read().split("\n") + [""]:
                if not line.strip():
                    if "processor" in cur:
                        core_of[int(cur["processor"])] = (cur.get("physical id", "0"), cur.get("core id",
                                                                                            cur["processor"]))
                    cur = {}
                    continue
                k, _, v = line.partition(":")
                cur[k.strip()] = v.strip()
                if k.strip() == "model name":
                    model = v.strip()
    except OSError:
        pass
    physical = len({core_of.get(c, c) for c in logical})
    return len(logical), max(1, physical), model


def cpu_baseline(w, runs=5):
    """The reference's own AVX2 8-bit kernels with its 16/64-bit cascade
    (oracle/_ref, compiled from the reference's sources) on the host cores,
    driven as src/algo/manager.c:120-185 drives them: T threads pulling
    1,000-sequence chunks, per-thread heaps merged in thread order.  T = the
    logical CPUs of this process (the reference's default, get_nprocs(),
    src/util/thread_pool.c:41-43, within the host's CPU share), the physical
    cores among them, and 1; each the median of `runs` runs over a sample of
    the DB sized to about a second per run (the whole DB at T = logical)."""
    so = os.path.join(ROOT, "oracle", "_ref", "libssaref.so")
    if not os.path.exists(so) or w["kind"] != "sw" or w["bw"] != 8:
        return None, None
    ref = C.CDLL(so)
    ref.ref_bench_sw8.restype = C.c_double
    ref.ref_bench_sw8.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_char_p, C.c_size_t,
                                  C.c_size_t, C.c_size_t, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    ref.ref_matrix_builtin.argtypes = [C.c_char_p]
    ref.ref_set_gaps.argtypes = [C.c_int, C.c_int]
    ref.ref_matrix_builtin(w["matrix"].encode())
    ref.ref_set_gaps(*w["gaps"])
    lut = np.zeros(256, np.uint8)
    for i, ch in enumerate("-ABCDEFGHIKLMNPQRSTVWXYZU*OJ"):
        if ch != "-":
            lut[ord(ch)] = i
            lut[ord(ch.lower())] = i
    db = w["db"]
    lens_all = np.array([len(s) for s in db], np.int64)
    offs_all = np.concatenate([[0], np.cumsum(lens_all)[:-1]]).astype(np.int64)
    codes = lut[np.frombuffer(b"".join(db), np.uint8)]
    qcodes = [lut[np.frombuffer(q, np.uint8)].tobytes() for q in w["queries"]]
    logical, physical, model = host_cpus()
    k = w["k"]

    def run(threads, n):
        """one pass of every query over the first n sequences: (GCUPS, seconds, top-k per query)"""
        cells, secs, tops = 0, 0.0, []
        for qc in qcodes:
            ids = np.zeros(k, np.int64)
            sc = np.zeros(k, np.int64)
            cnt = C.c_size_t()
            secs += ref.ref_bench_sw8(codes.ctypes.data, offs_all.ctypes.data, lens_all.ctypes.data, n, qc,
                                      len(qc), k, 1000, threads, ids.ctypes.data, sc.ctypes.data, C.byref(cnt))
            cells += len(qc) * int(lens_all[:n].sum())
            tops.append(list(zip(sc[:cnt.value].tolist(), ids[:cnt.value].tolist())))
        return cells / secs / 1e9, secs, tops

    table, check = {}, None
    for threads in sorted({logical, physical, 1}, reverse=True):
        n = len(db) if threads == logical else max(1000, len(db) * threads // max(logical, 16))
        n = min(n, len(db))
        vals, secs = [], []
        for r in range(runs):
            g, t, tops = run(threads, n)
            vals.append(g)
            secs.append(t)
            if threads == logical and r == 0:
                check = (n, tops)
        table["T%d" % threads] = {"threads": threads, "gcups_median": round(float(np.median(vals)), 2),
                                  "gcups_runs": [round(v, 2) for v in vals], "sequences": n,
                                  "seconds_per_run": round(float(np.median(secs)), 3)}
    best = table["T%d" % logical]
    return dict(value=best["gcups_median"], unit="GCUPS", cores=logical, kind="reference",
                threads=logical, physical_cores=physical, cpu_model=model,
                sample="reference AVX2 8-bit kernels with the 16/64-bit cascade (oracle/_ref), chunk 1000, per-thread "
                       "heaps merged in thread order; all %d queries; median of %d runs per thread count; T=%d over "
                       "the whole DB, fewer threads over the first sequences (about a second per run)"
                       % (len(qcodes), runs, logical),
                by_threads=table), check


def copy_probe(mib):
    """One hipMemcpy device-to-device of a known size: the reference point that
    turns the FETCH_SIZE / WRITE_SIZE counters of a profile into bytes.  Run the
    profile with HSA_ENABLE_SDMA=0, so that the copy is a kernel the counters
    see rather than a DMA-engine transfer."""
    hip = C.CDLL("libamdhip64.so")
    n = C.c_size_t(mib << 20)
    a, b = C.c_void_p(), C.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(C.byref(p), n) != 0:
            raise SystemExit("hipMalloc failed")
    hip.hipMemset(a, 1, n)
    hip.hipDeviceSynchronize()
    hip.hipMemcpy(b, a, n, 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)


def setup_protein(db):
    import libssa_amd as ssa
    from libssa_amd import synthetic as S
    ssa.reset_settings()
    ssa.set_output_mode(ssa.OUTPUT_ERROR)
    ssa.init_symbol_translation(ssa.AMINOACID, ssa.FORWARD_STRAND, 1, 1)
    ssa.init_score_matrix(ssa.MATRIX_BUILDIN, ssa.BLOSUM62)
    ssa.init_gap_penalties(-10, -1)
    tmp = tempfile.mkdtemp(prefix="libssa_bench_")
    path = os.path.join(tmp, "db.fasta")
    S.write_fasta(path, db)
    ssa.init_db(path)
    return path, tmp


def hits_key(hits):
    return [(a.score, a.db_id, a.alignment) for a in hits]


def bench_c4(args):
    import libssa_amd as ssa
    from libssa_amd import synthetic as S
    qs, db, _ = S.config_c4(scale=args.scale)
    path, tmp = setup_protein(db)
    hs = [ssa.init_sequence_fasta(ssa.READ_FROM_STRING, q.decode()) for q in qs]
    reqs = [(ssa.SEARCH_SW, h, 20, 8, ssa.COMPUTE_SCORE) for h in hs]
    tuned, t_tune, choices = [], 0.0, []
    if not args.no_autotune:
        ta = time.perf_counter()
        tuned = ssa.autotune(ssa.SEARCH_SW, [30, 60, 120, 250, 500, 1000, 2000], copies=4)
        t_tune = time.perf_counter() - ta
        choices = ssa.autotune_choices()
    ssa.batch_align(reqs)  # warm-up: builds the DB's streams for every lane width used
    t0 = time.perf_counter()
    plain = [ssa.sw_align(h, 20, 8, ssa.COMPUTE_SCORE) for h in hs]
    t1 = time.perf_counter()
    lists, stats = ssa.batch_align(reqs)
    t2 = time.perf_counter()
    cells = sum(st.cells for st in stats)
    equal = all(hits_key(a) == hits_key(b) for a, b in zip(plain, lists))
    os.remove(path)
    os.rmdir(tmp)
    return {
        "metric": "batched SW GCUPS on 1 MI355X (vs one query per launch)",
        "value": round(cells / (t2 - t1) / 1e9, 2), "unit": "GCUPS", "n_gpus": 1, "higher_is_better": True,
        "dtype": "fp16" if stats[0].cell_arithmetic == 2 else "int16", "data": "synthetic",
        "config": {"workload": "256 queries log-uniform 30-2000 vs 500k lognormal protein DB, 2% planted, "
                               "sw_align(q, 20, 8, scores), BLOSUM62 10/1", "db_sequences": len(db),
                   "queries": len(qs), "query_residues": sum(len(q) for q in qs)},
        "batched": {"seconds": round(t2 - t1, 3), "gcups": round(cells / (t2 - t1) / 1e9, 2),
                    "gpu_first_tier_ms": round(stats[0].ms_gpu_first, 3),
                    "graph_setup_ms": round(stats[0].ms_graph, 3), "graph_reused": bool(stats[0].graph_reused),
                    "d2h_bytes": int(sum(st.d2h_bytes for st in stats)),
                    "all_pairs_bytes": int(sum(12 * st.seq_count for st in stats))},
        "one_query_per_launch": {"seconds": round(t1 - t0, 3), "gcups": round(cells / (t1 - t0) / 1e9, 2)},
        "speedup": round((t1 - t0) / (t2 - t1), 3), "results_equal": equal, "cells": cells,
        "autotune": {"seconds": round(t_tune, 2),
                     "shapes": [{"qlen": L, "lanes_per_stream": g, "rows_per_lane": r, "columns_per_step": c[2],
                                 "profile_placement": c[3], "ms": round(ms, 3)}
                                for (L, g, r, ms), c in zip(tuned, choices)]},
    }


def serve_trace(ssa, reqs, handles, slots, max_batch, batch_cells=0.0):
    """Replays an arrival trace against a server; returns per-request
    (latency_ms, queue_ms, cells) and the wall time from first submit to last result."""
    import threading
    srv = ssa.Server(slots=slots, max_batch=max_batch)
    if batch_cells:
        srv.tune(batch_cells=batch_cells)
    tickets = [None] * len(reqs)
    t0 = time.perf_counter()

    def submitter():
        for i, r in enumerate(reqs):
            d = t0 + r["t"] - time.perf_counter()
            if d > 0:
                time.sleep(d)
            kind = ssa.SEARCH_SW if r["kind"] == "sw" else ssa.SEARCH_NW
            tickets[i] = srv.submit(kind, handles[i], 32, 8, ssa.COMPUTE_ALIGNMENT)

    th = threading.Thread(target=submitter)
    th.start()
    out = []
    for i in range(len(reqs)):
        while tickets[i] is None:
            time.sleep(0.0005)
        hits, st, lat, qms = srv.wait(tickets[i])
        out.append((lat, qms, st.cells, hits_key(hits)))
    th.join()
    wall = time.perf_counter() - t0
    c = srv.counters()
    srv.stop()
    return out, wall, c


def bench_c5(args):
    import numpy as np
    import libssa_amd as ssa
    from libssa_amd import synthetic as S
    db = S.config_c5(scale=args.scale)
    path, tmp = setup_protein(db)
    warm = [ssa.init_sequence_fasta(ssa.READ_FROM_STRING, q.decode()) for q in
            (S.make_query(1, 120), S.make_query(2, 700), S.make_query(3, 3000))]
    for h in warm:  # builds the DB's streams for the lane widths the mix uses
        ssa.sw_align(h, 32, 8, ssa.COMPUTE_ALIGNMENT)
        ssa.nw_align(h, 32, 8, ssa.COMPUTE_ALIGNMENT)
    rates = [float(x) for x in args.rates.split(",")]
    rows = []
    for rate in rates:
        reqs = S.c5_requests(args.requests, rate)
        handles = [ssa.init_sequence_fasta(ssa.READ_FROM_STRING, r["query"].decode()) for r in reqs]
        row = {"rate_per_s": rate, "requests": len(reqs)}
        results = {}
        for name, slots, mb in (("one_request_per_launch", 1, 1), ("continuous_batching", args.slots, 16)):
            out, wall, c = serve_trace(ssa, reqs, handles, slots, mb, args.batch_cells)
            lat = np.array([o[0] for o in out])
            cells = sum(o[2] for o in out)
            results[name] = [o[3] for o in out]
            row[name] = {"p50_ms": round(float(np.percentile(lat, 50)), 1),
                         "p99_ms": round(float(np.percentile(lat, 99)), 1),
                         "mean_queue_ms": round(float(np.mean([o[1] for o in out])), 1),
                         "throughput_per_s": round(len(out) / wall, 2), "gcups": round(cells / wall / 1e9, 1),
                         "batches": c.batches, "max_batch": c.max_batch_seen, "deferrals": c.deferrals}
        row["results_equal"] = results["one_request_per_launch"] == results["continuous_batching"]
        rows.append(row)
        for h in handles:
            ssa.free_sequence(h)
        print(json.dumps(row), file=sys.stderr, flush=True)
    os.remove(path)
    os.rmdir(tmp)
    return {
        "metric": "served p50/p99 latency and GCUPS on 1 MI355X (vs one request per launch)",
        "unit": "ms", "n_gpus": 1, "higher_is_better": False, "dtype": "fp16/int16", "data": "synthetic",
        "config": {"workload": "1M lognormal protein DB resident; Poisson arrivals; 80% SW / 20% NW, top 32 "
                               "with CIGARs, 8-bit start; |q| 70% U(50,400) + 30% log-U(400,5000)",
                   "db_sequences": len(db), "slots": args.slots, "batch_cells": args.batch_cells or 2.5e11},
        "rates": rows,
    }


def bench_dbload(args):
    """init_db + first search on the FASTA file and on its packed file."""
    import libssa_amd as ssa
    from libssa_amd import synthetic as S
    db = S.config_c5(args.scale)
    path, tmp = setup_protein(db)
    packed = os.path.join(tmp, "db.ssapack")
    q = ssa.init_sequence_fasta(ssa.READ_FROM_STRING, S.make_query(0x5EED0D0, 500).decode())

    def timed(fn):
        t0 = time.perf_counter()
        r = fn()
        return r, (time.perf_counter() - t0) * 1e3

    def load(p):
        _, t_init = timed(lambda: ssa.init_db(p))
        h1, t_first = timed(lambda: ssa.sw_align(q, 20, 8, ssa.COMPUTE_ALIGNMENT))
        h2, t_next = timed(lambda: ssa.sw_align(q, 20, 8, ssa.COMPUTE_ALIGNMENT))
        assert hits_key(h1) == hits_key(h2)
        return hits_key(h1), {"init_db_ms": round(t_init, 1), "first_search_ms": round(t_first, 1),
                              "next_search_ms": round(t_next, 1)}

    ssa.sw_align(q, 20, 8, ssa.COMPUTE_SCORE)  # GPU context, kernels loaded
    want, fasta = load(path)
    _, t_pack = timed(lambda: ssa.pack_db(packed))
    got, pk = load(packed)
    ssa.free_sequence(q)
    size = os.path.getsize(packed)
    for f in (path, packed):
        os.remove(f)
    os.rmdir(tmp)
    return {
        "metric": "init_db + first search latency on 1 MI355X (FASTA vs packed file)",
        "value": round(pk["init_db_ms"] + pk["first_search_ms"], 1), "unit": "ms", "n_gpus": 1,
        "higher_is_better": False, "dtype": "fp16", "data": "synthetic",
        "config": {"workload": "1M lognormal protein DB (median 300); sw_align 8-bit, |q| = 500, top 20 with CIGARs",
                   "db_sequences": len(db), "packed_file_bytes": size},
        "fasta": fasta, "packed": pk, "pack_ms": round(t_pack, 1), "results_equal": got == want,
    }


def bench_search(args, w, steps, warmup, cpu=False):
    """One search workload (c1, c2, c3): `steps` timed steps of every query
    through sw_align / nw_align with the DB resident.  Returns the JSON record."""
    import libssa_amd as ssa
    from libssa_amd import synthetic as S
    ssa.reset_settings()
    ssa.set_output_mode(ssa.OUTPUT_ERROR)
    ssa.init_symbol_translation(w["symtype"], w["strands"], 1, 1)
    if "const" in w:
        ssa.init_constant_scores(*w["const"])
    else:
        ssa.init_score_matrix(ssa.MATRIX_BUILDIN, w["matrix"])
    ssa.init_gap_penalties(*w["gaps"])
    tmp = tempfile.mkdtemp(prefix="libssa_bench_")
    path = os.path.join(tmp, "db.fasta")
    S.write_fasta(path, w["db"])
    ssa.init_db(path)
    qh = [ssa.init_sequence_fasta(ssa.READ_FROM_STRING, q.decode()) for q in w["queries"]]
    align = ssa.COMPUTE_ALIGNMENT if w.get("align") else ssa.COMPUTE_SCORE
    ssa.set_kernel_shape(*[int(x) for x in getattr(args, "kernel_shape", "0,0").split(",")])
    ssa.set_kernel_variant(*[int(x) for x in getattr(args, "kernel_variant", "0,0").split(",")])
    ssa.set_cell_arithmetic(getattr(args, "cell_arithmetic", 0))
    search = ssa.sw_align if w["kind"] == "sw" else ssa.nw_align

    def step():
        st = []
        tops = []
        for h in qh:
            hits = search(h, w["k"], w["bw"], align)
            st.append(ssa.last_stats())
            tops.append([(a.score, a.db_id) for a in hits])
        return st, tops

    for _ in range(warmup):
        step()
    t0 = time.perf_counter()
    stats = []
    tops = []
    for _ in range(steps):
        st, tops = step()
        stats.extend(st)
    t1 = time.perf_counter()
    trace = None
    if getattr(args, "task_trace", False):
        # one more search per query with the first tier's tasks recorded
        trace = {}
        ssa.set_task_trace(True)
        try:
            for q, h in zip(w["queries"], qh):
                search(h, w["k"], w["bw"], align)
                trace["q%d" % len(q)] = task_trace_summary(ssa.task_trace())
        finally:
            ssa.set_task_trace(False)
    for h in qh:
        ssa.free_sequence(h)
    os.remove(path)
    os.rmdir(tmp)
    ssa.set_kernel_shape(0, 0)
    ssa.set_kernel_variant(0, 0)
    ssa.set_cell_arithmetic(0)
    # the first tier's kernel of each query in the last step
    kernels = [{"lanes_per_stream": s.lanes_per_stream, "rows_per_lane": s.rows_per_lane,
                "columns_per_step": s.columns_per_step, "profile_placement": s.profile_placement,
                "cell_arithmetic": "fp16" if s.cell_arithmetic == 2 else "int16"} for s in stats[-len(qh):]]

    secs = t1 - t0
    cells_step = sum(s.cells for s in stats) / steps
    # NT both strands: the DB holds each sequence and its reverse complement, and
    # both query strands run against both, as in the reference; the unique work
    # is half of that
    dup = 2 if w.get("strands") == 3 else 1
    gcups = cells_step / dup * steps / secs / 1e9
    kernel_ms = sum(s.ms_gpu_first for s in stats)
    # the roofline counts the cells the first-tier kernels scored: a
    # both-strand search reads its reverse strand from the forward strand's
    # reverse-complement twins where that is exact, and scores half
    scored_step = sum(s.cells_scored for s in stats) / steps
    ops = sum(s.cells_scored for s in stats) * OPS_PER_CELL[w["kind"]]
    achieved = ops / (kernel_ms / 1e3) / 1e12
    traffic = measured_traffic(args.workload if args.workload != "all" else "c2") if cpu else None
    out = {
        "metric": "SW/NW GCUPS on 1 MI355X",
        "value": round(gcups, 2),
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": round(secs / steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        # the first tier's cell arithmetic: packed fp16 holding integer scores
        # exactly (per-lane windows for SW), or packed saturating int16
        "dtype": "fp16" if any(k["cell_arithmetic"] == "fp16" for k in kernels) else "int16",
        "data": "synthetic",
        "config": {"workload": w["desc"], "db_sequences": len(w["db"]), "cells_per_step": int(cells_step / dup),
                   "reference_cells_per_step": int(cells_step), "scored_cells_per_step": int(scored_step),
                   "query_lengths": [len(q) for q in w["queries"]], "hits": w["k"], "bit_width": w["bw"],
                   "kernels": kernels},
        "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": round(PEAK_VALU_TOPS, 1),
                     "unit": "TOP/s", "frac": round(achieved / PEAK_VALU_TOPS, 4),
                     "traffic": traffic,
                     "traffic_source": "committed rocprofv3 FETCH_SIZE + WRITE_SIZE of this workload "
                                       "(profiles/score_kernel_hbm_traffic.txt), bytes per step" if traffic else None,
                     "kernel": "score_kernel, the first-tier launch of each search (HIP events on the library's stream)",
                     "ops_per_cell": OPS_PER_CELL[w["kind"]],
                     "kernel_ms_per_step": round(kernel_ms / steps, 3),
                     "launches_per_step": len(w["queries"])},
        "breakdown_ms_per_step": {
            "gpu_first_tier": round(kernel_ms / steps, 3),
            "gpu_requeue": round(sum(s.ms_gpu_requeue for s in stats) / steps, 3),
            "host_replay": round(sum(s.ms_replay for s in stats) / steps, 3),
            "traceback": round(sum(s.ms_traceback for s in stats) / steps, 3)},
        "host_transfer": {"d2h_bytes_per_step": int(sum(s.d2h_bytes for s in stats) / steps),
                          # what the host would copy if it decoded every pair: 12 bytes per
                          # (scored query buffer, entry) pair, one scored buffer per search here
                          "all_pairs_bytes_per_step": int(sum(12 * s.seq_count for s in stats) / steps),
                          "replayed_chunks_per_step": round(sum(s.replayed_chunks for s in stats) / steps, 1),
                          "chunks_per_search": stats[0].chunk_count if stats else 0},
        "cpu_baseline": None,
    }
    if trace is not None:
        out["task_trace"] = trace
    if cpu and not args.no_cpu_baseline:
        cb, check = cpu_baseline(w)
        if cb is not None and check[0] == len(w["db"]):
            # the reference ran on the whole DB: its top-k per query against
            # this run's (scores as a multiset; the hits above the k-th score
            # by ID; at the k-th score the reference's threads pick among ties
            # in an order that depends on their timing)
            same_scores, same_above = True, True
            for mine, ref in zip(tops, check[1]):
                same_scores &= sorted(x[0] for x in mine) == sorted(x[0] for x in ref)
                kth = min(x[0] for x in ref) if ref else None
                same_above &= sorted(x for x in mine if x[0] > kth) == sorted(x for x in ref if x[0] > kth)
            cb["agrees_with_gpu"] = {"queries": len(tops), "top_k_scores_equal": same_scores,
                                     "hits_above_kth_score_equal": same_above}
        out["cpu_baseline"] = cb
    return out


def task_trace_summary(recs):
    """Statistics of a traced first-tier launch (ssa.task_trace records, in
    task order).  Each workgroup runs its first task by its index, so the
    first `grid` records are the first round, and workgroup w holds slot
    w // CUs of its CU.  `us_per_step_by_slot` is the first round's time per
    4-column step of a wave in each slot (all slots busy); `ms_at_slot_rates`
    is the launch's task steps at the per-CU rate those give, with no tail."""
    if not recs:
        return None
    import libssa_amd as ssa
    a = np.array([r[:4] for r in recs], dtype=np.float64)
    start, end, wg, steps = a.T
    dur = end - start
    grid = int(wg.max()) + 1
    cus = max(1, ssa.device_cus())
    first = (np.arange(len(recs)) < grid) & (steps > 0)
    slot = (wg // cus).astype(np.int64)
    us = []
    for k in range(int(slot.max()) + 1):
        m = first & (slot == k)
        us.append(round(float((dur[m] / steps[m]).mean()) / 1e3, 3) if m.any() else None)
    rate = sum(1 / u for u in us if u)  # steps per us per CU
    wg_end = np.zeros(grid)
    np.maximum.at(wg_end, wg.astype(np.int64), end)
    lt = int(np.argmax(steps))
    return {"tasks": len(recs), "workgroups": grid, "cus": cus, "launch_ms": round(float(end.max()) / 1e6, 3),
            "task_steps": int(steps.sum()), "us_per_step_by_slot": us,
            "ms_at_slot_rates": round(float(steps.sum()) / (cus * rate) / 1e3, 3) if rate else None,
            "workgroup_end_ms": {p: round(float(np.percentile(wg_end, q)) / 1e6, 3)
                                 for p, q in (("p10", 10), ("p50", 50), ("p90", 90), ("max", 100))},
            "longest_task": {"steps": int(steps[lt]), "workgroup": int(wg[lt]),
                             "start_ms": round(float(start[lt]) / 1e6, 3), "end_ms": round(float(end[lt]) / 1e6, 3)}}


def summary(rec, keys):
    return {k: rec[k] for k in keys if k in rec}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="all",
                    help="all (default: c2 headline plus c3, c4, c5), or one of c1 c2 c3 c4 c5 dbload")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--scale", type=float, default=1.0, help="c4/c5: DB size factor")
    ap.add_argument("--rates", default="5,10,20,30", help="c5: arrival rates per second")
    ap.add_argument("--no-autotune", action="store_true", help="c4: keep the cost model's shapes")
    ap.add_argument("--slots", type=int, default=2, help="c5: server workers for continuous batching")
    ap.add_argument("--batch-cells", type=float, default=0.0, help="c5: cell budget per batch (0 = default)")
    ap.add_argument("--requests", type=int, default=60, help="c5: requests per rate")
    ap.add_argument("--kernel-shape", default="0,0", metavar="G,R",
                    help="c1-c3: force lanes per stream pair and rows per lane (0,0 = chosen per query)")
    ap.add_argument("--kernel-variant", default="0,0", metavar="C,P",
                    help="c1-c3: force columns per step (4, 8) and profile placement (1 LDS, 2 VGPRs); 0 = chosen")
    ap.add_argument("--cell-arithmetic", type=int, default=0, choices=[0, 1],
                    help="c1-c3: 0 = fp16 where it holds (default), 1 = packed int16 everywhere")
    ap.add_argument("--task-trace", action="store_true",
                    help="c1-c3: also trace one more search per query and summarise its score launch's tasks")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="also run one device-to-device copy of MIB MiB (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (--gpus %d requested)" % args.gpus)
        sys.exit(2)

    if args.workload in ("c4", "c5", "dbload"):
        fn = {"c4": bench_c4, "c5": bench_c5, "dbload": bench_dbload}[args.workload]
        print(json.dumps(fn(args)))
        return
    if args.workload in ("c1", "c2", "c3"):
        out = bench_search(args, workload(args.workload), args.steps, args.warmup, cpu=args.workload == "c2")
        if args.copy_probe:
            copy_probe(args.copy_probe)
        print(json.dumps(out))
        return
    if args.workload != "all":
        raise SystemExit("unknown workload " + args.workload)

    # the headline: the SW 8-bit cascade workload (value, roofline, CPU baseline)
    def note(msg):
        print("bench: " + msg, file=sys.stderr, flush=True)

    note("c2 headline: SW 8-bit cascade, 200k DB, q = 50/500/5000")
    out = bench_search(args, workload("c2"), args.steps, args.warmup, cpu=True)
    keys = ("value", "unit", "ms_per_step", "config", "roofline", "breakdown_ms_per_step")
    # the same workload with the cell arithmetic forced to packed int16: the
    # default fp16 datapath holds only integers of at most 2,047 in magnitude
    # (SW: relative to per-lane offsets), which fp16 represents and adds
    # exactly, so both give the same bit-exact results
    note("c2 in packed int16")
    i16_args = argparse.Namespace(**vars(args))
    i16_args.cell_arithmetic = 1
    i16 = bench_search(i16_args, workload("c2"), min(args.steps, 10), 2)
    out["int16_arithmetic"] = dict(summary(i16, ("value", "unit", "ms_per_step", "dtype")),
                                   kernel_ms_per_step=i16["roofline"]["kernel_ms_per_step"],
                                   frac=i16["roofline"]["frac"],
                                   note="the headline workload with ssa_amd_set_cell_arithmetic(1): every cell in "
                                        "packed saturating int16; the default computes the same integers in fp16, "
                                        "exactly (bit-identical results)")
    # the metric's other parts, each on its own workload at full size
    note("c3: NW DNA both strands with CIGARs, 1M DB")
    out["nw_dna"] = summary(bench_search(args, workload("c3"), 10, 2), keys)
    note("c4: 256 queries batched vs one query per launch, 500k DB")
    out["batched"] = bench_c4(args)
    note("c5: served SW/NW mix, 1M DB, one request per launch vs continuous batching")
    c5_args = argparse.Namespace(**vars(args))
    c5_args.rates = "10,30"
    out["served"] = bench_c5(c5_args)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
