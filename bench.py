"""Smith-Waterman database search benchmark on one MI355X.

Default workload: the full-size plain search -- 5M database sequences (about
1.75G residues, log-normal lengths, mean 350, tail to 35,000), queries of
length 144, 464, 1000, 2005 and 5478 searched one query per call, BLOSUM62,
gap 10/2, exact int32 scores for every sequence and the top-20 hits.
A step is one pass of all five queries over the resident database.

Prints ONE JSON line (metric GCUPS, higher is better).  Extra keys:
  roofline      the dominant scan kernel (k_scan) against MI355X's VALU issue
                peak: VALU instructions issued per second (the committed
                SQ_INSTS_VALU count per cell pair x the cell pairs of the
                timed launches / their HIP-event time) over 256 CUs x 64
                lane-instructions x 2.4 GHz; beside it the 9-op rate of the
                reference's recurrence and the HBM side, as flat fields
  cpu_baseline  this build's restatement of the reference's CPU path (the
                host AVX2 scan of libswhybrid_host.so: int8 pass, then int16
                for lanes that may have saturated, over a packed database) on
                every 3rd sequence of the database with all five queries,
                timed on this host's cores (its CPU share: OMP_NUM_THREADS
                where set), with the host's model and core counts and the
                scan's committed calibration against the reference's own
                kernels
  per_query     GCUPS and kernel time for each query length
  batched, batched_small_db, dbfile, hybrid, served, overflow_stress
                the other workloads (64-query batch, small database, database
                files through the worker pipeline, CPU + GPU workers, the
                serving loop, 2M sequences with near-copies), each skippable
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# MI355X: 256 CUs x 2.4 GHz x 64 lanes x 2 (packed 16-bit) lane-ops per clock
PEAK_INT16_TOPS = 256 * 2.4e9 * 64 * 2 / 1e12
# VALU issue: one wave64 instruction per SIMD every 4 clocks = 64 lane-instructions per CU clock
PEAK_VALU_TIPS = 256 * 2.4e9 * 64 / 1e12
OPS_PER_CELL = 9   # 1 add, 3 sub, 5 max (the reference's ONE_CELL_UPDATE)
# Fewest VALU instructions a packed cell pair needs with this instruction set:
# 1 add, 3 sub, 2 max, 1 three-way max, and half a three-way max of the row's
# running-maximum tree.  Issue at this count is the work the recurrence needs.
MIN_INSTS_PER_CELL_PAIR = 7.5
PEAK_HBM_GBS = 8000.0
TOP_K = 20


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_cpu():
    """Model name and core counts of this host."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "nproc": os.cpu_count(),
            "affinity": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None}


def cpu_threads():
    """This process's CPU share: OMP_NUM_THREADS where the host sets it, else
    every core it may run on."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    return host_cpu()["affinity"] or os.cpu_count() or 1


def cpu_calibration():
    path = os.path.join(ROOT, "profiles", "cpu_baseline_calibration.json")
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def cpu_baseline(db, queries, matrix, go, ge, stride=3):
    """The reference's CPU path restated (the host AVX2 scan, int8 then int16)
    over one sequence in `stride` of the same database, every query, on this
    host's cores; the database is packed first (untimed, as the GPU leg's is
    resident)."""
    from swhybrid_amd.datagen import SeqDB
    from swhybrid_amd.engine import cpu_search
    idx = np.arange(0, db.n, stride)
    sub = SeqDB(db.residues, db.offsets[idx].copy(), db.lengths[idx].copy())
    threads = cpu_threads()
    cells = 0
    t = pack = 0.0
    for q in queries:
        scores, secs, psecs = cpu_search(sub, q, matrix=matrix, gap_open=go, gap_extend=ge,
                                         threads=threads)
        if (scores < 0).any():
            raise RuntimeError("cpu baseline: a score beyond int16 on this workload")
        t += secs
        pack += psecs
        cells += sub.total * len(q)
    out = {
        "value": round(cells / t / 1e9, 2), "unit": "GCUPS", "cores": threads, "kind": "port",
        "sample": (f"one sequence in {stride} of the same database ({sub.n} sequences, "
                   f"{sub.total} residues) x query lengths {[len(q) for q in queries]}; "
                   f"this build's AVX2 restatement of the reference's avxCompute + "
                   f"avxCompute16 (int8 pass, int16 for saturated lanes) over the packed "
                   f"sample; {t:.1f} s scan, {pack:.1f} s packing (not in value)"),
        "host": host_cpu(),
    }
    cal = cpu_calibration()
    if cal:
        out["calibration"] = cal
    return out


def cpu_sweep():
    """Query segmentation of the host scan: GCUPS per segment length at query
    lengths 5478 and 9000 over 30k sequences of the 5M database's length
    distribution, on this host's cores and on one."""
    from swhybrid_amd import datagen as G
    from swhybrid_amd.engine import cpu_search
    db = G.c2_db(n=30_000)
    res = {"host": host_cpu(), "workload": f"{db.n} sequences, {db.total} residues",
           "gcups": {}}
    for threads in sorted({1, cpu_threads()}):
        for L in (5478, 9000):
            q = G.queries([L], 12)[0]
            ref = None
            for seg in (-1, 4096, 2048, 1024, 512):
                best = 1e30
                for _ in range(2):
                    got, secs, _ = cpu_search(db, q, threads=threads, query_segment=seg)
                    best = min(best, secs)
                if ref is None:
                    ref = got
                elif not np.array_equal(got, ref):
                    raise RuntimeError(f"segment {seg} changed a score at query length {L}")
                key = f"threads={threads} qlen={L} segment={'whole' if seg < 0 else seg}"
                res["gcups"][key] = round(db.total * L / best / 1e9, 1)
                log(f"{key}: {res['gcups'][key]} GCUPS")
    print(json.dumps(res))


# Algorithmic bytes per k_scan launch of this workload: the packed database
# once (1 byte per residue, lane blocks padded: 1.82 GB) plus the 20 MB score
# vector and the profile.
ALGORITHMIC_BYTES_PER_LAUNCH = 1.84e9


def measured_counters():
    """HBM bytes per k_scan launch and VALU instructions per cell pair,
    averaged over the five query lengths of this workload, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE (calibrated on a 1 GiB copy) and
    SQ_INSTS_VALU runs; the counters cannot be read from inside a timed run."""
    path = os.path.join(ROOT, "profiles", "k_scan_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
        return t.get("mean_per_launch_bytes"), t.get("mean_valu_insts_per_cell_pair")
    except (OSError, ValueError):
        return None, None


TUNE_BUCKETS = (96, 192, 384, 768, 1536, 3072, 5000)


def run_batched(eng, db, d, plain_gcups, group_min_len):
    """64 queries, lengths log-uniform in [50, 5000], one call: one persistent
    launch whose device work queue holds every (query, block group) item.
    The autotuner first picks, per query-length bucket, the launch shape on a
    1/8 length-stratified sample; the database is re-laid out if it prefers
    another group-mode length; the call is then captured in a hipGraph and
    the timed call replays it.  Returns (result, database handle to keep)."""
    from swhybrid_amd import datagen as G
    tune = eng.autotune(db, TUNE_BUCKETS, sample_every=8, queries_per_bucket=4)
    log(f"autotune: {tune['seconds']:.1f} s, group_min_len {tune['group_min_len']}, "
        f"{[(t['qlen_max'], t['strip_width'], t['profile_mode'], t['group_lanes'], t['group_width']) for t in tune['entries']]}")
    want_min = tune["group_min_len"]
    # the default layout's threshold grows with the database (at least 2048,
    # 2,672 here): a tuned 2048 keeps it, a longer or shorter one re-lays out
    cur_min = group_min_len if group_min_len > 0 else 2048
    if want_min != cur_min:
        d.close()
        t0 = time.perf_counter()
        d = eng.load_db(db, group_min_len=-1 if want_min >= 2**31 else want_min)
        log(f"database re-laid out for group_min_len {want_min} in {time.perf_counter() - t0:.1f} s")
    qs = G.c3_queries()
    # every score of every query comes back: into one page-locked buffer
    # (64 x 5M x 4 B), which the device copy writes directly
    out = eng.pinned_scores(db.n, len(qs))
    t0 = time.perf_counter()
    _, st0 = eng.search_batch(d, qs, use_graph=1, out=out)     # eager run, then capture
    first = time.perf_counter() - t0
    t0 = time.perf_counter()
    scores, st = eng.search_batch(d, qs, use_graph=1, out=out)   # graph replay
    wall = time.perf_counter() - t0
    del scores, out
    log(f"batched: {len(qs)} queries, {st['cells']:.3e} cells in {wall:.2f} s "
        f"(first call {first:.2f} s)")
    g = st["cells"] / wall / 1e9
    return {"workload": "64 queries, lengths log-uniform [50, 5000], one call, all scores "
                        "returned", "queries": len(qs), "sum_qlen": int(sum(len(q) for q in qs)),
            "gcups": round(g, 1),
            "kernel_gcups": round(st["cells"] / (st["kernel_ms"] / 1e3) / 1e9, 1),
            "wall_s": round(wall, 3), "first_call_s": round(first, 3),
            "eager_kernel_gcups": round(st0["cells"] / (st0["kernel_ms"] / 1e3) / 1e9, 1),
            "vs_plain": round(g / plain_gcups, 3), "n_overflow": st["n_overflow"],
            "graph": "replayed (profile upload, scan, rescore, copy-back in one hipGraph)",
            "autotune": {"seconds": round(tune["seconds"], 1),
                         "sample": f"every 8th sequence by length ({tune['sample_seqs']} "
                                   f"sequences, {tune['sample_residues']} residues), "
                                   "4 queries per bucket",
                         "group_min_len": want_min,
                         "buckets": [{k: (round(v, 6) if isinstance(v, float) else v)
                                      for k, v in t.items() if k != "lane_cost"}
                                     for t in tune["entries"]]}}, d


def run_small_db(eng):
    """The small database: 100k sequences (29.9M residues) are too few for one
    query's lane items to fill 256 CUs (780 lane blocks, 195 work items
    against 512-768 workgroup slots), so the database also carries a layout
    with its sequences split over lane groups, which launches of a few
    queries use; one batched launch of the 64 queries has lane items enough
    and runs on the lane layout."""
    from swhybrid_amd import datagen as G
    db = G.c1_db()
    d = eng.load_db(db)
    qs = G.c3_queries()
    for q in qs[:4]:
        eng.search(d, q)
    t0 = time.perf_counter()
    cells = 0
    for q in qs:
        _, st = eng.search(d, q)
        cells += st["cells"]
    plain = time.perf_counter() - t0
    eng.search_batch(d, qs, use_graph=1)                # eager run, then capture
    batched = 1e30
    for _ in range(3):                                  # graph replays, best of 3
        t0 = time.perf_counter()
        _, st = eng.search_batch(d, qs, use_graph=1)
        batched = min(batched, time.perf_counter() - t0)
    q1 = G.c1_query()
    for _ in range(3):
        eng.search(d, q1)
    one = 1e30
    for _ in range(5):                                  # a 2 ms call: best of 5
        t0 = time.perf_counter()
        _, st1 = eng.search(d, q1)
        one = min(one, time.perf_counter() - t0)
    d.close()
    log(f"small database: plain {cells / plain / 1e9:.0f} GCUPS, batched "
        f"{st['cells'] / batched / 1e9:.0f} GCUPS, its own query {st1['cells'] / one / 1e9:.0f} GCUPS")
    return {"workload": "100k sequences (29.9M residues, the CPU-runnable database) x the 64 "
                        "batched queries: one call per query vs one batched call; and its "
                        "own query (length 256) alone",
            "plain_gcups": round(cells / plain / 1e9, 1),
            "batched_gcups": round(st["cells"] / batched / 1e9, 1),
            "batched_vs_plain": round(plain / batched, 2),
            "c1_query_gcups": round(st1["cells"] / one / 1e9, 1),
            "c1_query_ms": round(one * 1e3, 2)}


def run_overflow_stress():
    """The overflow-heavy workload: 2M sequences, 10% of them 3,001-35,000
    residues long, 20% replaced by ~90%-identity near-copies of the two
    queries (3000, 5478), both queries in one batched call, with BLOSUM62
    10/2 and BLOSUM50 13/2.  BLOSUM50 near-copies of the 5478 query reach the
    int16 pass's exact limit and are rescored in int32 on the device."""
    from swhybrid_amd import datagen as G
    from swhybrid_amd.engine import Engine
    t0 = time.perf_counter()
    db, qs, planted = G.c5_db()
    log(f"overflow stress: {db.n} sequences, {db.total} residues ({time.perf_counter() - t0:.0f} s)")
    out = {"workload": "2M sequences (10% of 3,001-35,000 residues, 20% near-copies of the "
                       "queries), queries 3000 + 5478 in one batched call",
           "nseq": db.n, "residues": db.total}
    for matrix, go, ge in (("blosum62", 10, 2), ("blosum50", 13, 2)):
        e = Engine(matrix, go, ge)
        d = e.load_db(db)
        e.search_batch(d, qs)
        t0 = time.perf_counter()
        scores, st = e.search_batch(d, qs)
        wall = time.perf_counter() - t0
        out[matrix] = {"gcups": round(st["cells"] / wall / 1e9, 1),
                       "wall_s": round(wall, 3), "n_overflow": st["n_overflow"],
                       "group_cells_frac": round(st["cells_group"] / st["cells"], 3),
                       "max_score": int(scores.max()),
                       "copies_over_10k": int((scores[:, planted] > 10_000).sum())}
        log(f"overflow stress {matrix} {go}/{ge}: {out[matrix]['gcups']} GCUPS, "
            f"{st['n_overflow']} rescored in int32")
        d.close()
        e.close()
    return out


def run_dbfile(eng):
    """The reference's database files: the 100k-sequence database written as
    FASTA, made into .seq/.map/.title, mapped, and searched through the worker
    pipeline (DataPool over the file's batches -> HIPWorker), the reference's
    own path for one query, with the top-10 checked against the resident
    search of the same sequences."""
    import tempfile
    from swhybrid_amd import alphabet as A
    from swhybrid_amd import datagen as G
    from swhybrid_amd import dbfile as F
    from swhybrid_amd.engine import topk
    db = G.c1_db()
    q = G.c1_query()
    letters = np.frombuffer(A.CODE_ORDER.encode() + b"@", dtype=np.uint8)
    with tempfile.TemporaryDirectory() as tmp:
        fa = os.path.join(tmp, "c1.fa")
        with open(fa, "wb") as f:
            text = letters[db.residues]
            for i in range(db.n):
                o = int(db.offsets[i])
                f.write(b">seq%d\n" % i + text[o:o + int(db.lengths[i])].tobytes() + b"\n")
        fa_bytes = os.path.getsize(fa)
        t0 = time.perf_counter()
        records, slots = F.make(fa, os.path.join(tmp, "c1"))
        make_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        f = F.DbFile(os.path.join(tmp, "c1"))
        fdb = f.seqdb()
        open_s = time.perf_counter() - t0
        f.pool_search(q, eng.matrix, k=10)                       # warm
        r = f.pool_search(q, eng.matrix, k=10)
        d = eng.load_db(fdb)
        scores, _ = eng.search(d, q)
        d.close()
        ts, ti = topk(scores, 10)
        same = bool((ts == r["top_score"]).all() and (ti == r["top_index"]).all())
        # each hit's title names the generated sequence whose residues it holds
        titles_ok = all(np.array_equal(fdb.seq(int(i)), db.seq(int(f.title(int(i))[4:])))
                        for i in ti)
        f.close()
    log(f"database files: make {fa_bytes / make_s / 1e6:.0f} MB/s, open {open_s * 1e3:.0f} ms, "
        f"pipeline {r['gcups']:.0f} GCUPS")
    return {"workload": "100k-sequence database (29.9M residues) as FASTA -> .seq/.map/.title -> "
                        "worker pipeline over the file's batches, query length 256, top-10",
            "fasta_mb": round(fa_bytes / 1e6, 1), "make_s": round(make_s, 2),
            "make_mb_per_s": round(fa_bytes / make_s / 1e6, 1), "open_ms": round(open_s * 1e3, 1),
            "records": records, "slots": slots,
            "pipeline_gcups": round(r["gcups"], 1), "pipeline_s": round(r["seconds"], 3),
            "top10_equals_resident_search": same, "titles_ok": titles_ok}


def pipeline_breakdown(db, qlen, chunk_bytes=0, runs=2, first_chunk_bytes=0):
    """The worker pipeline (DataPool -> HIPWorker, no CPU worker) over the
    5M-sequence database in host memory for one query length: where the
    HIPWorker's time goes (first fill, aux fills and their swh_db_create
    part, searches and their kernels, waits for the aux thread), from the
    last of `runs` calls."""
    from swhybrid_amd import datagen as G
    from swhybrid_amd.engine import hybrid_search
    q = G.queries([qlen], 5000 + qlen)[0]
    for _ in range(runs):
        r = hybrid_search(db, q, k=TOP_K, hip_workers=1, cpu_threads=0, chunk_bytes=chunk_bytes,
                          first_chunk_bytes=first_chunk_bytes)
    rep = r["report"]
    cells = db.total * qlen
    out = {"qlen": qlen, "chunk_bytes": rep["chunk_bytes"],
           "first_chunk_bytes": rep["first_chunk_bytes"],
           "gcups_reference_measure": round(rep["gcups"], 1),
           "gcups_wall": round(cells / rep["wall_seconds"] / 1e9, 1),
           "seconds": round(rep["seconds"], 4), "wall_seconds": round(rep["wall_seconds"], 4)}
    for k in ("dev_first_fill_s", "dev_aux_fill_s", "dev_pack_s", "dev_search_s",
              "dev_aux_wait_s", "dev_kernel_s"):
        out[k] = round(rep[k], 4)
    out["dev_chunks"] = rep["dev_chunks"]
    out["kernel_gcups"] = round(cells / rep["dev_kernel_s"] / 1e9, 1) if rep["dev_kernel_s"] else None
    log(f"pipeline qlen {qlen}: {json.dumps(out)}")
    return out


def run_hybrid(db, qlens=(1000, 5478)):
    """The reference's hybrid scheduling: the worker pipeline over the same
    5M-sequence database held in host memory, chunks of the pool's default
    schedule (512 MB below query length 600, 256 MB below 2,500, from there a
    256 MB first chunk, then 1 GB) to the HIPWorker and of 4 MB to the CPU
    worker (the host's AVX2 cores), all
    from one DataPool; HIPWorker alone first, then both.  `*_gcups` is the
    reference's own measure (first worker start to last stop, after the
    first chunk is filled); `*_wall_gcups` times the whole call."""
    from swhybrid_amd import datagen as G
    from swhybrid_amd.engine import hybrid_search
    # the HIPWorker's own host work (chunk copies, upload) competes with the
    # CPU worker's threads: 4 threads is where the CPU costs it least
    threads = min(4, os.cpu_count() or 1)
    out = {"workload": "worker pipeline over the 5M-sequence database in host memory: "
                       "HIPWorker alone vs HIPWorker + CPU worker on one DataPool",
           "cpu_threads": threads}
    for L in qlens:
        q = G.queries([L], 5000 + L)[0]
        a = hybrid_search(db, q, k=TOP_K, hip_workers=1, cpu_threads=0)
        b = hybrid_search(db, q, k=TOP_K, hip_workers=1, cpu_threads=threads)
        same = bool((a["top_score"] == b["top_score"]).all())
        ra, rb = a["report"], b["report"]
        cells = db.total * L
        out[str(L)] = {"hip_only_gcups": round(ra["gcups"], 1), "hybrid_gcups": round(rb["gcups"], 1),
                       "hip_only_wall_gcups": round(cells / ra["wall_seconds"] / 1e9, 1),
                       "hybrid_wall_gcups": round(cells / rb["wall_seconds"] / 1e9, 1),
                       "cpu_share": round(rb["cpu_residues"] / max(1, rb["cpu_residues"] + rb["hip_residues"]), 4),
                       "cpu_worker_gcups": round(rb["cpu_residues"] * L / rb["cpu_busy_s"] / 1e9, 1)
                       if rb["cpu_busy_s"] > 0 else 0.0,
                       "top20_scores_equal": same}
        log(f"hybrid qlen {L}: HIP alone {ra['gcups']:.0f} ({out[str(L)]['hip_only_wall_gcups']:.0f} "
            f"wall), HIP + CPU {rb['gcups']:.0f} ({out[str(L)]['hybrid_wall_gcups']:.0f} wall) GCUPS "
            f"(CPU share {out[str(L)]['cpu_share']:.3f})")
    return out


def run_serving(eng, d, qps_levels, n_requests, round_ms=0.0, slices=0, wait_weight=0.0,
                queue="auto"):
    """Poisson arrivals at fixed QPS, lengths logN(mean 400, sigma 0.7) clipped
    to [30, 5000], top-20 per request, the database resident; p50/p99 latency
    (submit to top-k delivered) and sustained GCUPS per level.  Each request
    is also timed alone first (one search call, the device otherwise idle):
    no scheduler can deliver it faster, so `slowdown` = latency / alone time
    is what the serving loop adds (waiting and sharing the device)."""
    import threading
    from swhybrid_amd import datagen as G
    from swhybrid_amd.engine import Server
    res = {}
    # burst: the 64 batched queries submitted at once (throughput of the loop
    # at saturation, next to the one batched call)
    qs = G.c3_queries()
    with Server(eng, d, round_ms=round_ms, slices=slices, wait_weight=wait_weight,
                queue=queue) as srv:
        t0 = time.perf_counter()
        ids = [srv.submit(q, k=TOP_K) for q in qs]
        got = [srv.wait(i) for i in ids]
        wall = time.perf_counter() - t0
        st = srv.stats()
    cells = sum(r["info"]["cells"] for r in got)
    res["burst"] = {"workload": "the 64 batched queries submitted at once", "requests": len(qs),
                    "gcups": round(cells / wall / 1e9, 1), "wall_s": round(wall, 3),
                    "rounds": st["rounds"], "queue": queue,
                    "first_done_ms": round(min(r["info"]["latency_ms"] for r in got), 1)}
    log(f"serving burst: {len(qs)} requests in {wall:.2f} s = {res['burst']['gcups']} GCUPS "
        f"({st['rounds']} rounds)")
    alone_ms = {}
    for qps in qps_levels:
        arrivals, qs = G.c4_requests(n_requests, qps)
        for q in qs:
            key = q.tobytes()
            if key not in alone_ms:
                eng.search(d, q)
                t0 = time.perf_counter()
                eng.search(d, q)
                alone_ms[key] = (time.perf_counter() - t0) * 1e3
        srv = Server(eng, d, round_ms=round_ms, slices=slices, wait_weight=wait_weight,
                     queue=queue)
        ids = [None] * len(qs)
        t0 = time.perf_counter()

        def feed():
            for i, (a, q) in enumerate(zip(arrivals, qs)):
                delay = t0 + a - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
                ids[i] = srv.submit(q, k=TOP_K)

        th = threading.Thread(target=feed)
        th.start()
        lat, done_ms, first_submit = [], [], None
        slow, rounds = [], []
        cells = 0
        for i in range(len(qs)):
            while ids[i] is None:
                time.sleep(0.0005)
            r = srv.wait(ids[i])
            info = r["info"]
            lat.append(info["latency_ms"])
            slow.append(info["latency_ms"] / alone_ms[qs[i].tobytes()])
            rounds.append(info["rounds"])
            done_ms.append(info["done_ms"])
            first_submit = info["submit_ms"] if first_submit is None else min(first_submit, info["submit_ms"])
            cells += info["cells"]
        th.join()
        st = srv.stats()
        srv.close()
        span = (max(done_ms) - first_submit) / 1e3
        lat = np.array(lat)
        slow = np.array(slow)
        alone = np.array([alone_ms[q.tobytes()] for q in qs])
        res[str(int(qps)) if float(qps).is_integer() else str(qps)] = {
            "requests": len(qs), "p50_ms": round(float(np.percentile(lat, 50)), 1),
            "p99_ms": round(float(np.percentile(lat, 99)), 1),
            "mean_ms": round(float(lat.mean()), 1),
            "alone_p50_ms": round(float(np.percentile(alone, 50)), 1),
            "alone_p99_ms": round(float(np.percentile(alone, 99)), 1),
            "slowdown_p50": round(float(np.percentile(slow, 50)), 2),
            "slowdown_p99": round(float(np.percentile(slow, 99)), 2),
            "gcups": round(cells / span / 1e9, 1), "offered_tcups": round(
                qps * np.mean([len(q) for q in qs]) * d.residues / 1e12, 2),
            "peak_inflight": st["peak_inflight"], "rounds": st["rounds"],
            "max_rounds_per_request": int(max(rounds)), "queue": queue}
        r_ = res[list(res)[-1]]
        log(f"serving {qps} QPS: p50 {r_['p50_ms']} ms, p99 {r_['p99_ms']} ms (alone "
            f"{r_['alone_p50_ms']} / {r_['alone_p99_ms']}), slowdown p50 {r_['slowdown_p50']} "
            f"p99 {r_['slowdown_p99']}, {r_['gcups']} GCUPS, {r_['rounds']} rounds")
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nseq", type=int, default=5_000_000, help="database sequences")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--group-min-len", type=int, default=0)
    ap.add_argument("--query-lengths", type=str, default="",
                    help="comma-separated subset of the workload's query lengths (profiling runs)")
    ap.add_argument("--no-batched", action="store_true", help="skip the 64-query batched run")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving run")
    ap.add_argument("--no-dbfile", action="store_true", help="skip the database-file run")
    ap.add_argument("--no-hybrid", action="store_true", help="skip the CPU + GPU pipeline run")
    ap.add_argument("--no-overflow-stress", action="store_true",
                    help="skip the 2M-sequence near-copy workload")
    ap.add_argument("--serve-qps", type=str, default="5,20,50")
    ap.add_argument("--serve-queue", choices=["auto", "live", "rounds"], default="auto",
                    help="serving loop: live queue (the default on this database) or rounds")
    ap.add_argument("--serve-requests", type=int, default=60)
    ap.add_argument("--serve-round-ms", type=float, default=0.0,
                    help="device work per serving round in ms (0 = the library default)")
    ap.add_argument("--serve-wait-weight", type=float, default=0.0,
                    help="priority credit of waiting (0 = the library default)")
    ap.add_argument("--serve-slices", type=int, default=0,
                    help="0 = budgeted rounds, n = every request in n fixed slices")
    ap.add_argument("--pipeline", type=int, default=0, metavar="QLEN",
                    help="only run the worker pipeline at this query length and print where "
                         "its time goes")
    ap.add_argument("--pipeline-chunk-mb", type=int, default=0)
    ap.add_argument("--pipeline-first-mb", type=int, default=0,
                    help="first chunk of the schedule (later chunks double up to the chunk size)")
    ap.add_argument("--cpu-sweep", action="store_true",
                    help="only measure the host scan's query segmentation and print it")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.cpu_sweep:
        cpu_sweep()
        return
    if args.copy_calibration:
        import torch
        n = args.copy_calibration << 20
        src = torch.empty(n, dtype=torch.uint8, device="cuda:0")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": n}))
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: --gpus {args.gpus} is not supported "
              "(the search shards by database block; see the design notes)", file=sys.stderr)
        sys.exit(2)

    from swhybrid_amd import datagen as G
    from swhybrid_amd.engine import Engine, topk

    t0 = time.perf_counter()
    db = G.c2_db(n=args.nseq)
    if args.pipeline:
        print(json.dumps(pipeline_breakdown(db, args.pipeline, args.pipeline_chunk_mb << 20,
                                            first_chunk_bytes=args.pipeline_first_mb << 20)))
        return
    queries = G.c2_queries()
    if args.query_lengths:
        keep = {int(x) for x in args.query_lengths.split(",")}
        queries = [q for q in queries if len(q) in keep]
    log(f"database: {db.n} sequences, {db.total} residues, max len {int(db.lengths.max())} "
        f"({time.perf_counter() - t0:.1f} s)")
    eng = Engine("blosum62", 10, 2, device=0)
    t0 = time.perf_counter()
    d = eng.load_db(db, group_min_len=args.group_min_len)
    log(f"packed + uploaded in {time.perf_counter() - t0:.1f} s: {d.info()}")

    for w in range(args.warmup):
        for q in queries:
            eng.search(d, q)
        log(f"warmup step {w + 1}/{args.warmup} done")

    # the caller keeps one page-locked score vector: the device writes it
    # directly, and the top-k is taken from it before the next call
    out = eng.pinned_scores(db.n)
    cells = 0
    kernel_ms = 0.0
    per_query = {}
    tops = {}
    t_start = time.perf_counter()
    for s in range(args.steps):
        for q in queries:
            scores, st = eng.search(d, q, out=out)
            tops[len(q)] = topk(scores, TOP_K)
            cells += st["cells"]
            kernel_ms += st["kernel_ms"]
            pq = per_query.setdefault(len(q), {"cells": 0, "kernel_ms": 0.0, "total_ms": 0.0})
            pq["cells"] += st["cells"]
            pq["kernel_ms"] += st["kernel_ms"]
            pq["total_ms"] += st["total_ms"]
        log(f"step {s + 1}/{args.steps} done")
    elapsed = time.perf_counter() - t_start
    gcups = cells / elapsed / 1e9

    # dominant kernel: k_scan (one launch per call: long pairs, then lane blocks).
    # Its bound is VALU issue: `achieved` is VALU lane-instructions issued per
    # second, the committed SQ_INSTS_VALU count per packed cell pair x the cell
    # pairs of the timed launches / their HIP-event time, against one wave64
    # instruction per SIMD every 4 clocks at 2.4 GHz.
    secs = kernel_ms / 1e3
    traffic, insts_per_pair = measured_counters()
    # the committed counters were read on the full workload; a run on another
    # database or a subset of the queries reports no figures derived from them
    full = db.n == 5_000_000 and not args.query_lengths
    if not full:
        traffic = insts_per_pair = None
    ops9 = OPS_PER_CELL * cells / secs / 1e12 if secs > 0 else 0.0
    tips = cells / 2 * insts_per_pair / secs / 1e12 if (insts_per_pair and secs > 0) else None
    useful = cells / 2 * MIN_INSTS_PER_CELL_PAIR / secs / 1e12 if secs > 0 else 0.0
    avg_ms = kernel_ms / (args.steps * len(queries))
    roofline = {
        "bound": "valu", "kernel": "k_scan",
        "achieved": round(tips, 3) if tips else None, "peak": round(PEAK_VALU_TIPS, 2),
        "unit": "T VALU lane-instructions/s",
        "frac": round(tips / PEAK_VALU_TIPS, 4) if tips else None,
        "traffic": traffic,
        "valu_insts_per_cell_pair": insts_per_pair,
        # the same launches at the fewest instructions the recurrence needs
        # (7.5 per cell pair): the share of the issue peak doing that work
        "useful_achieved": round(useful, 3),
        "useful_frac": round(useful / PEAK_VALU_TIPS, 4),
        # the reference's 9 ops per cell against 2 packed 16-bit lane-ops per
        # lane and clock: above 1 because one v_pk_minimum3_f16 does two of
        # the nine maxima, so it is not a fraction of a hardware limit
        "ops9_achieved": round(ops9, 3), "ops9_peak": round(PEAK_INT16_TOPS, 1),
        "ops9_unit": "T packed int16 lane-ops/s (9 per cell)",
        "ops9_ratio": round(ops9 / PEAK_INT16_TOPS, 4),
        "kernel_gcups": round(cells / secs / 1e9, 1) if secs > 0 else None,
        "avg_launch_ms": round(avg_ms, 2),
    }
    if traffic:
        # the HBM side of the same launches: measured bytes / launch time
        gbs = traffic / (avg_ms / 1e3) / 1e9
        roofline.update({"hbm_achieved": round(gbs, 1), "hbm_peak": PEAK_HBM_GBS,
                         "hbm_unit": "GB/s", "hbm_frac": round(gbs / PEAK_HBM_GBS, 4),
                         "hbm_algorithmic_bytes": ALGORITHMIC_BYTES_PER_LAUNCH})
    pq_out = {str(k): {"gcups": round(v["cells"] / (v["total_ms"] / 1e3) / 1e9, 1),
                       "kernel_gcups": round(v["cells"] / (v["kernel_ms"] / 1e3) / 1e9, 1),
                       "kernel_ms": round(v["kernel_ms"] / args.steps, 2)}
              for k, v in per_query.items()}
    out = {
        # the headline metric; `value` is its plain per-query leg on the
        # full-size database, the batched and served legs and p50/p99 follow
        "metric": "GCUPS on 1 MI355X (batched/served vs plain per-query) + p50/p99 latency",
        "value": round(gcups, 1), "unit": "GCUPS", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int16x2 saturating scan + int32 rescore",
        "data": "synthetic (seeded generator, Robinson-Robinson residue frequencies)",
        "config": {"workload": (f"plain search, {db.n:,} sequences ({db.total / 1e9:.2f}G "
                                f"residues), query lengths "
                                f"{'/'.join(str(len(q)) for q in queries)}, BLOSUM62 10/2, "
                                f"top-20"),
                   "nseq": db.n, "residues": db.total, "query_lengths": [len(q) for q in queries]},
        "roofline": roofline,
        "per_query": pq_out,
        "top1": {str(k): int(v[0][0]) for k, v in tops.items()},
    }
    if not args.no_batched:
        out["batched"], d = run_batched(eng, db, d, gcups, args.group_min_len)
        out["batched_small_db"] = run_small_db(eng)
    if not args.no_dbfile:
        out["dbfile"] = run_dbfile(eng)
    if not args.no_hybrid:
        out["hybrid"] = run_hybrid(db)
    if not args.no_serve:
        out["served"] = run_serving(eng, d, [float(x) for x in args.serve_qps.split(",")],
                                    args.serve_requests, args.serve_round_ms,
                                    args.serve_slices, args.serve_wait_weight,
                                    args.serve_queue)
    d.close()
    eng.close()
    if not args.no_overflow_stress:
        out["overflow_stress"] = run_overflow_stress()
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(db, queries, eng.matrix, 10, 2)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
