#!/usr/bin/env python3
"""Benchmark of the read-classification path on one MI355X.

Headline workload (the largest single-GPU configuration): 200,000 references
of ~8 Gbp (log-normal lengths, 30 % of them holding 1-3 segments of 1-20 kb
from a pool of 1,000 shared segments), k = 20 (T = 38), 10,000,000 reads x
150 bp with 1 % substitutions, seeded synthetic data.  A step classifies all
10M reads in one batched call (both strands of every read: probe, filters,
chains, scoring, best hit, strand merge) with the reads already in HBM.  The
same index is then served (continuous batching of a Poisson request stream)
and classified in the plain one-launch-per-batch form.

Extra keys: workload 2 (5,000 references, k = 16, 1M x 150 bp), the
mixed-length batch (2M reads of 75 bp-20 kb) against its index, the served
stream of workload 4 (20,000 references, 2 Gbp), .mhm load and FASTA parse.
Prints one JSON line; progress goes to stderr.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TIER_NAMES = ("second chance", "middle", "HBM workspaces", "last resort")


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def device_tensors(torch, qb, qo):
    d_reads = torch.from_numpy(qb).to("cuda")
    d_off = torch.from_numpy(qo.view(np.int64)).to("cuda")
    return d_reads, d_off


def classify_checked(ix, native, *args, **kw):
    """A classify call made again when the library grew its tie pool."""
    for attempt in range(4):
        try:
            return ix.classify_device(*args, **kw)
        except native.MirageError as e:
            if e.status != native.MH_ERR_OVERFLOW or attempt == 3:
                raise


def run_steps(torch, handle, ix, d_reads, d_off, n, flags, steps, warmup, native, top_per_read=16):
    """Warm-up, then `steps` timed classify calls bracketed by synchronisation.
    Returns (seconds per step, per-step kernel timings, counters, per-kernel
    counters, results, top pool)."""
    d_out = torch.empty(n * native.READ_RESULT_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
    top_cap = top_per_read * n + 1024
    d_top = torch.empty(top_cap * 4, dtype=torch.uint8, device="cuda")
    for _ in range(max(1, warmup)):
        tn, _ = classify_checked(ix, native, d_reads, d_off, n, d_out, d_top, top_cap, flags=flags)
        assert tn <= top_cap
    torch.cuda.synchronize()
    timings = []
    stream = torch.cuda.ExternalStream(handle.stream_ptr())
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    timed = not (flags & native.MH_CLASSIFY_GRAPH)
    for _ in range(steps):
        ix.classify_device(d_reads, d_off, n, d_out, d_top, top_cap,
                           flags=flags | (native.MH_CLASSIFY_TIMING if timed else 0), wait=False,
                           sync_inputs=False)
        if timed:
            timings.append(handle.timing())
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ev_s = ev0.elapsed_time(ev1) / 1000.0
    counters = handle.counters()
    kcounters = handle.kernel_counters()
    out = d_out.cpu().numpy().view(native.READ_RESULT_DTYPE).copy()
    top = d_top.cpu().numpy().view(np.uint32).copy()
    del d_out, d_top
    return max(wall, ev_s) / steps, timings, counters, kcounters, out, top


def same_results(native, a_out, a_top, b_out, b_top):
    return bool(np.array_equal(a_out["highest"].view(np.uint32), b_out["highest"].view(np.uint32))
                and np.array_equal(a_out["strand"], b_out["strand"])
                and native.tie_sets(a_out, a_top) == native.tie_sets(b_out, b_top))


def kernel_names(tiles):
    """Row names of mh_get_kernel_counters for a tile table."""
    names = []
    for i in range(8):
        if i < len(tiles):
            ml, b, e, _ = tiles[i]
            team = f"LaneTeam<{b}>" if b < 64 else f"BlockTeam<{b}>"
            names.append(f"k_classify<{team}> tile {i} (reads <= {ml} bp, {e} entries)")
        else:
            names.append(f"length tile {i} (unused)")
    return names + [f"overflow tier: {t}" for t in TIER_NAMES]


def kernel_table(timings, kcounters, tiles, tiers, read_len):
    """Per kernel: mean time per call (HIP events on the library stream), the
    strands it finished and their algorithmic bytes: a 16-byte directory
    pair per probe, 8 bytes per bucket entry read, 2 bytes per extension
    comparison, the read's bases and a 32-byte result per strand.
    `read_len` is the bases per strand: one number, or one per row (the 8
    length tiles, then the 4 overflow tiers)."""
    rows = []
    names = kernel_names(tiles)
    tier_kernels = [f"k_classify<BlockTeam<{tiers[0][1]}>> (second chance)",
                    f"k_classify<BlockTeam<{tiers[1][1]}>> (middle tier)",
                    "k_classify_hbm<256> (HBM workspaces)", "k_classify_hbm<1024> (last resort)"]
    for k in range(12):
        if k < 8:
            if k >= len(tiles):
                continue
            ms = float(np.mean([t["bucket_ms"][k] for t in timings]))
            name = names[k]
        else:
            ms = float(np.mean([t["tier_ms"][k - 8] for t in timings]))
            name = tier_kernels[k - 8]
        c = kcounters[k]
        rl = read_len[k] if isinstance(read_len, (list, tuple)) else read_len
        alg = (16 * c["probes"] + 8 * c["entries"] + 2 * c["ext_compares"] +
               (rl + 32) * c["queries"])
        rows.append({"kernel": name, "ms": round(ms, 4), "strands": c["queries"],
                     "probes": c["probes"], "entries": c["entries"], "cands": c["cands"],
                     "ext_compares": c["ext_compares"], "alg_bytes": int(alg),
                     "alg_GBs": round(alg / max(ms, 1e-6) / 1e6, 2)})
    return rows


def profiled_traffic(fname, kernel):
    """(HBM bytes per launch, source) from a committed profiler summary of
    `kernel` on the full-size workload, or (None, None)."""
    path = os.path.join(ROOT, "profiles", fname)
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        tj = json.load(f)
    if tj.get("kernel") != kernel:
        return None, None
    return int(tj["traffic_bytes"]), tj.get("source")


def roofline_of(row, traffic=None, traffic_src=None):
    achieved = row["alg_bytes"] / (row["ms"] / 1e3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
            "traffic_source": traffic_src, "kernel": row["kernel"], "kernel_ms": row["ms"],
            "alg_bytes_per_launch": row["alg_bytes"], "strands": row["strands"]}


def make_requests(rng, qb, qo, n_requests):
    """Requests of floor(10^U(0,4)) consecutive reads of the pool."""
    sizes = np.floor(10 ** rng.uniform(0, 4, n_requests)).astype(int)
    starts = rng.integers(0, len(qo) - 1 - sizes)
    return [(qb[int(qo[a]):int(qo[a + n])], qo[a:a + n + 1] - qo[a]) for a, n in
            zip(starts, sizes)], sizes


def serve_stream(native, ix, reqs, sizes, loads, rng, log, max_batch_reads=65536, max_inflight=3,
                 flags=0):
    """Saturation (all requests submitted at once) first, then the given
    fractions of it as Poisson offered loads; latency is submit -> results
    on the host, measured by the server.  The client submits from arrays
    prepared beforehand and a waiter thread collects results into
    preallocated buffers, so that the client is not what limits the rate."""
    import threading
    reqs = [(np.ascontiguousarray(b, dtype=np.uint8), np.ascontiguousarray(o, dtype=np.uint64))
            for b, o in reqs]
    nmax = int(max(sizes))

    def run(rate):
        srv = native.Server(ix, hbm_budget=16 << 30, max_inflight=max_inflight,
                            max_batch_reads=max_batch_reads, flags=flags)
        ids, lat = [], []
        out = np.zeros(nmax, native.READ_RESULT_DTYPE)
        top = np.zeros(64 * nmax + 1024, np.uint32)

        def waiter():
            for i in range(len(reqs)):
                while i >= len(ids):
                    time.sleep(0.0001)
                need, inf = srv.wait_into(ids[i], out, top)  # raises on a failed request
                assert need <= len(top)
                lat.append((inf["done_s"] - inf["submit_s"], inf["submit_s"], inf["done_s"],
                            inf["admit_s"] - inf["submit_s"], inf["done_s"] - inf["admit_s"]))

        th = threading.Thread(target=waiter)
        th.start()
        gaps = rng.exponential(1.0 / rate, len(reqs)) if rate else np.zeros(len(reqs))
        nxt = t_sub = time.perf_counter()
        for (b, o), g in zip(reqs, gaps):
            if rate:
                nxt += g
                d = nxt - time.perf_counter()
                if d > 0:
                    time.sleep(d)
            ids.append(srv.submit_arrays(b, o))
        t_sub = time.perf_counter() - t_sub
        th.join()
        srv.stop()
        lt = np.array([x[0] for x in lat])
        wt = np.array([x[3] for x in lat])
        st = np.array([x[4] for x in lat])
        span = max(x[2] for x in lat) - min(x[1] for x in lat)
        # where the latency goes: waiting for a slot and HBM budget (submit
        # -> admit), then the batch's copies, kernels and hand-back (admit ->
        # done); at the 99th percentile request, its own split
        i99 = int(np.argsort(lt)[min(len(lt) - 1, int(np.ceil(0.99 * len(lt))) - 1)])
        return {"reads_per_s": float(sizes.sum() / span), "span_ms": span * 1e3,
                "submit_ms": t_sub * 1e3,
                "p50_ms": float(np.percentile(lt, 50) * 1e3),
                "p99_ms": float(np.percentile(lt, 99) * 1e3), "requests": len(reqs),
                "wait_p50_ms": float(np.percentile(wt, 50) * 1e3),
                "wait_p99_ms": float(np.percentile(wt, 99) * 1e3),
                "service_p50_ms": float(np.percentile(st, 50) * 1e3),
                "service_p99_ms": float(np.percentile(st, 99) * 1e3),
                "p99_request": {"reads": int(sizes[i99]), "wait_ms": float(wt[i99] * 1e3),
                                "service_ms": float(st[i99] * 1e3)}}

    run(0)  # warm-up: slots, streams, first launches
    sat = run(0)
    log(f"serving saturation: {sat['reads_per_s'] / 1e6:.2f} M reads/s, p50 {sat['p50_ms']:.1f} ms "
        f"({sat['span_ms']:.1f} ms, of which submitting {sat['submit_ms']:.1f} ms)")
    out = {"saturation": sat, "loads": [], "max_batch_reads": max_batch_reads,
           "max_inflight": max_inflight}
    mean_reads = float(sizes.mean())
    for f in loads:
        rate = f * sat["reads_per_s"] / mean_reads
        r = run(rate)
        r["offered_fraction"] = f
        r["offered_requests_per_s"] = rate
        out["loads"].append(r)
        log(f"serving load {f:.2f}: {r['reads_per_s'] / 1e6:.2f} M reads/s, p50 "
            f"{r['p50_ms']:.2f} ms, p99 {r['p99_ms']:.2f} ms (wait p99 {r['wait_p99_ms']:.2f}, "
            f"service p99 {r['service_p99_ms']:.2f} ms)")
    return out


def plain_stream(ix, reqs, sizes):
    """The plain form on the same request stream: one call per request, one
    launch per batch (no coalescing, no overlap), host reads in and results
    out as the server does.  Returns reads/s."""
    t = time.perf_counter()
    for b, o in reqs:
        ix.classify_host(b, o, flags=0)
    return float(sizes.sum() / (time.perf_counter() - t))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(oix, qb, qo, sample, threads, gpu_out, gpu_top, sample1, build_s, what):
    """The oracle (C++ restatement of the reference CPU path) on a bounded
    sample of the reads: one warm-up, then the median of 5 runs with
    `threads` threads, and the same on a smaller sample with 1 thread.  Also
    checks the GPU results on the sample."""
    from oracle import oracle as O
    from mirage_amd import native

    def timed(n, nt):
        # top_cap: room for the largest tie sets (segments shared by >100 references)
        sq, so = qb[: int(qo[n])], qo[: n + 1]
        oix.classify(sq, so, n_threads=nt, top_cap=1024)  # warm-up
        runs, res = [], None
        for _ in range(5):
            t0 = time.perf_counter()
            res = oix.classify(sq, so, n_threads=nt, top_cap=1024)
            runs.append(time.perf_counter() - t0)
        return float(np.median(runs)), res

    n = min(sample, len(qo) - 1)
    med, (out, top, _) = timed(n, threads)
    n1 = min(sample1, len(qo) - 1)
    med1, _ = timed(n1, 1)
    g = gpu_out[:n]
    same = (np.array_equal(g["highest"].view(np.uint32), out["highest"].view(np.uint32))
            and np.array_equal(g["strand"], out["strand"])
            and native.tie_sets(g, gpu_top) == O.top_sets(out, top))
    return {"value": round(n / med, 1), "unit": "reads/s", "cores": threads, "kind": "port",
            "sample": f"first {n} reads of the same workload, {what}; median of 5 after 1 "
                      f"warm-up (oracle index {build_s:.1f} s, not timed)",
            "bit_exact_vs_gpu": bool(same), "value_1thread": round(n1 / med1, 1),
            "sample_1thread": f"first {n1} reads, median of 5",
            "host_nproc": os.cpu_count(), "host_cpu": cpu_model()}


def mhm_leg(torch, native, handle, ix, info, rb, ro, threshold, d_reads, d_off, n, out, top,
            log):
    """Writes the workload's index as a .mhm v4 file (host memory file
    system when there is one) and loads it back onto the device three times:
    end-to-end load time from the mapped file, including the copy to HBM and
    the decode kernels.  The loaded index must equal the built one and
    classify the reads identically."""
    import shutil
    import tempfile
    d = "/dev/shm" if os.path.isdir("/dev/shm") else tempfile.gettempdir()
    need = 24 * info["n_entries"] + (2 << 30)
    if shutil.disk_usage(d).free < need:
        log(f"mhm: skipped, {d} has less than {need / 2**30:.1f} GiB free")
        return {"skipped": f"{d} too small"}
    path = os.path.join(d, f"mh_bench_{os.getpid()}.mhm")
    full = handle.build_index(rb, ro, threshold, last_bucket_empty=False)
    try:
        t = time.perf_counter()
        # 8-byte hashes and stream positions, the 64-bit build's format: the
        # file is larger than 4 GiB
        full.save_mhm(path, hash_bytes=8, streampos_bytes=8)
        save_s = time.perf_counter() - t
    finally:
        full.free()
    size = os.path.getsize(path)
    with native.MhmFile(path) as f:
        finfo = f.info()
    log(f"mhm: wrote {size / 2**30:.2f} GiB in {save_s:.2f} s ({finfo['n_buckets']} buckets)")
    loads, same_index, same = [], True, None
    keys = ("n_entries", "n_buckets", "empty_hash", "n_refs", "threshold")
    try:
        for rep in range(3):
            torch.cuda.synchronize()
            t = time.perf_counter()
            lx = handle.load_mhm(path)
            loads.append(time.perf_counter() - t)
            li = lx.info()
            same_index &= all(li[k] == info[k] for k in keys)
            if rep == 0:
                lout = torch.empty(n * native.READ_RESULT_DTYPE.itemsize, dtype=torch.uint8,
                                   device="cuda")
                ltop = torch.empty(len(top) * 4, dtype=torch.uint8, device="cuda")
                lx.classify_device(d_reads, d_off, n, lout, ltop, len(top),
                                   flags=native.MH_CLASSIFY_BATCHED)
                o2 = lout.cpu().numpy().view(native.READ_RESULT_DTYPE)
                t2 = ltop.cpu().numpy().view(np.uint32)
                same = same_results(native, o2, t2, out, top)
                del lout, ltop
            lx.free()
    finally:
        os.remove(path)
    best = min(loads)
    log(f"mhm: load {best:.3f} s ({size / best / 1e9:.2f} GB/s of file), index same "
        f"{same_index}, results same {same}")
    return {"file_bytes": size, "toc_entries": finfo["n_buckets"],
            "record_bytes": finfo["record_bytes"], "save_s": round(save_s, 3),
            "load_s": round(best, 4), "load_s_all": [round(x, 4) for x in loads],
            "load_file_GBps": round(size / best / 1e9, 3), "dir": d,
            "index_same_as_build": bool(same_index), "results_same": same}


def fasta_leg(torch, native, handle, qb, qo, log):
    """The reads as 61-column FASTA text in HBM, parsed on the device
    (mh_fasta_parse) five times: time per parse and GB/s of text; the parsed
    reads must equal the generated ones."""
    n = len(qo) - 1
    lens = np.diff(qo)
    if not np.all(lens == lens[0]):
        return {"skipped": "reads of one length only"}
    m = int(lens[0])
    seq = qb.reshape(n, m)
    parts = np.split(seq, list(range(61, m, 61)), axis=1)
    nl = np.full((n, 1), ord("\n"), np.uint8)
    body = np.concatenate([x for p in parts for x in (p, nl)], axis=1)
    text = b"".join(b">read%d\n" % i + body[i].tobytes() for i in range(n))
    d_text = torch.from_numpy(np.frombuffer(text, np.uint8).copy()).cuda()
    torch.cuda.synchronize()
    times = []
    for _ in range(5):
        t0 = time.perf_counter()
        pseq, pso, phdr, pho = handle.parse_fasta(d_text)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    same = bool(pseq.cpu().numpy().tobytes() == qb.tobytes() and pso.numel() == n + 1)
    best = min(times)
    log(f"fasta: {len(text) / 2**20:.0f} MiB of text, parse {best * 1e3:.2f} ms "
        f"({len(text) / best / 1e9:.1f} GB/s), reads equal {same}")
    return {"text_bytes": len(text), "records": n, "parse_ms": round(best * 1e3, 3),
            "parse_ms_all": [round(x * 1e3, 3) for x in times],
            "text_GBps": round(len(text) / best / 1e9, 2), "reads_equal": same}


def headline_leg(args, torch, native, D):
    """Workload 5 at full size: batched (eager and hipGraph replay), plain,
    served, per-kernel roofline, CPU oracle on a sample."""
    w = D.workload(5, args.w5_scale)
    w.reads.n = args.w5_reads
    t = time.perf_counter()
    rb, ro = D.make_refs(w.refs, w.seed)
    qb, qo, _ = D.make_reads(w.reads, rb, ro, w.seed)
    n = len(qo) - 1
    log(f"workload 5: {len(ro) - 1} refs, {len(rb) / 1e9:.2f} Gbp, {n} reads "
        f"({time.perf_counter() - t:.1f} s)")
    handle = native.Handle(0, args.max_read_len)
    t = time.perf_counter()
    ix = handle.build_index(rb, ro, w.threshold)
    torch.cuda.synchronize()
    index_s = time.perf_counter() - t
    info = ix.info()
    tiles, tiers = handle.get_tiles(ix)
    log(f"index: {info['n_entries']:,} entries, {info['device_bytes'] / 2**30:.1f} GiB, "
        f"{index_s:.2f} s; tiles {tiles[0]}, tiers {tiers}")
    d_reads, d_off = device_tensors(torch, qb, qo)
    res = {}
    # The batched call with and without the probe pass (MH_CLASSIFY_PREPROBE:
    # probe records for the first tile, which the overflow tiers also read);
    # the faster eager form is replayed as a hipGraph and gives the kernel
    # table.
    forms = {"batched": native.MH_CLASSIFY_BATCHED,
             "batched with the probe pass": native.MH_CLASSIFY_BATCHED | native.MH_CLASSIFY_PREPROBE}
    if args.w5_form == "batched":
        del forms["batched with the probe pass"]
    elif args.w5_form == "preprobe":
        del forms["batched"]
    eager = {}
    for name, fl in forms.items():
        eager[name] = run_steps(torch, handle, ix, d_reads, d_off, n, fl, args.steps, args.warmup,
                                native)
        log(f"{name} (eager): {eager[name][0] * 1e3:.2f} ms/step, "
            f"{n / eager[name][0] / 1e6:.2f} M reads/s")
    form = min(eager, key=lambda k: eager[k][0])
    sec, timings, ctr, kctr, out, top = eager[form]
    flags = forms[form]
    others = [k for k in eager if k != form]
    res["form"] = form
    res["eager_forms_ms_per_step"] = {k: round(v[0] * 1e3, 3) for k, v in eager.items()}
    res["forms_same_results"] = (same_results(native, eager[others[0]][4], eager[others[0]][5],
                                              out, top) if others else None)
    del eager
    gsec, _, _, _, gout, gtop = run_steps(torch, handle, ix, d_reads, d_off, n,
                                          flags | native.MH_CLASSIFY_GRAPH,
                                          args.steps, args.warmup, native)
    res["graph_replay_same_results"] = same_results(native, gout, gtop, out, top)
    log(f"{form} (hipGraph replay): {gsec * 1e3:.2f} ms/step, {n / gsec / 1e6:.2f} M reads/s, "
        f"same {res['graph_replay_same_results']}")
    res["eager_ms_per_step"] = round(sec * 1e3, 3)
    res["graph_ms_per_step"] = round(gsec * 1e3, 3)
    res["sec"] = min(sec, gsec)
    res["counters"] = ctr
    res["entries_per_probe"] = round(ctr["entries"] / max(ctr["probes"], 1), 2)
    res["tiles"], res["tiers"] = [list(x) for x in tiles], [list(x) for x in tiers]
    res["escalations"] = ctr["overflow"]
    res["call_ms"] = {"total": round(float(np.mean([t_["total_ms"] for t_ in timings])), 3),
                      "probe_pass": round(float(np.mean([t_["probe_ms"] for t_ in timings])), 3),
                      "merge": round(float(np.mean([t_["merge_ms"] for t_ in timings])), 3)}
    rows = kernel_table(timings, kctr, tiles, tiers, 150)
    res["kernels"] = rows
    dom = max(rows, key=lambda r: r["ms"])
    traffic, src = profiled_traffic("k_classify_workload5_traffic.json", dom["kernel"])
    res["roofline"] = roofline_of(dom, traffic, src)
    for r in rows:
        if r["strands"]:
            log(f"  {r['kernel']}: {r['ms']:.2f} ms, {r['strands']} strands, "
                f"{r['alg_GBs']} GB/s algorithmic")
    if not args.no_plain:
        psec, ptim, _, _, pout, ptop = run_steps(torch, handle, ix, d_reads, d_off, n, 0,
                                                 max(2, args.steps // 4), 1, native)
        res["plain"] = {"reads_per_s": round(n / psec, 1), "ms_per_step": round(psec * 1e3, 3),
                        "same_results": same_results(native, pout, ptop, out, top),
                        "batched_over_plain": round(psec / res["sec"], 3)}
        log(f"plain: {psec * 1e3:.1f} ms/step, {n / psec / 1e6:.2f} M reads/s")
    del d_reads, d_off
    torch.cuda.empty_cache()
    if not args.no_serve:
        rng = np.random.default_rng(w.seed)
        reqs, sizes = make_requests(rng, qb, qo, args.serve_requests)
        # the server runs the headline's form (with the probe pass when that
        # was the faster batched call)
        sflags = flags & native.MH_CLASSIFY_PREPROBE
        sv = serve_stream(native, ix, reqs, sizes, (0.25, 0.5, 0.9), rng, log,
                          max_batch_reads=args.serve_batch, max_inflight=args.serve_slots,
                          flags=sflags)
        sv["form"] = form
        sv["workload"] = (f"requests of floor(10^U(0,4)) reads (mean {sizes.mean():.0f}) drawn "
                          "from the 10M reads, Poisson arrivals; latency submit -> results on "
                          "the host")
        if not args.no_plain:
            sv["plain_reads_per_s"] = plain_stream(ix, reqs[: args.plain_requests],
                                                   sizes[: args.plain_requests])
            sv["served_over_plain"] = round(sv["saturation"]["reads_per_s"] /
                                            sv["plain_reads_per_s"], 2)
            log(f"plain on the request stream: {sv['plain_reads_per_s'] / 1e6:.2f} M reads/s")
        res["served"] = sv
    ix.free()
    handle.close()
    if not args.no_cpu:
        from oracle import oracle as O
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        ns = min(args.w5_cpu_sample, n)
        t = time.perf_counter()
        oix = O.Index.for_reads(rb, ro, w.threshold, qb[: int(qo[ns])], qo[: ns + 1])
        build_s = time.perf_counter() - t
        log(f"cpu baseline: oracle index of the sample's buckets in {build_s:.1f} s; "
            f"{ns} reads, {threads} threads")
        res["cpu_baseline"] = cpu_baseline(
            oix, qb, qo, ns, threads, out, top, max(1, ns // 10), build_s,
            "the oracle's index restricted to the buckets these reads probe (the same results "
            "as the full index for them)")
        log(f"cpu: {res['cpu_baseline']['value'] / 1e3:.1f} k reads/s, gpu bit-exact on the "
            f"sample: {res['cpu_baseline']['bit_exact_vs_gpu']}")
        del oix
    res.update({"n": n, "refs": len(ro) - 1, "ref_bases": int(ro[-1]),
                "index_entries": info["n_entries"], "index_GiB": round(info["device_bytes"] / 2**30, 2),
                "index_build_s": round(index_s, 2), "threshold": w.threshold})
    return res


def workload2_leg(args, torch, native, D):
    """Workload 2 (5,000 references, k = 16, 1M x 150 bp) and, on its index,
    the mixed-length batch of workload 3; .mhm and FASTA legs."""
    w = D.workload(2, 1.0)
    rb, ro, qb, qo, _ = D.generate(w)
    n = len(qo) - 1
    handle = native.Handle(0, 32768)
    ix = handle.build_index(rb, ro, w.threshold)
    info = ix.info()
    tiles, tiers = handle.get_tiles(ix)
    d_reads, d_off = device_tensors(torch, qb, qo)
    r = {"workload": "5000 refs lognormal(50 kb, 1.0) GC 40-60%, k=16 (T=30), "
                     f"{n} reads x 150 bp, 2% subs, 5% exact, 10% random"}
    sec, tim, ctr, kctr, out, top = run_steps(torch, handle, ix, d_reads, d_off, n,
                                              native.MH_CLASSIFY_BATCHED, args.steps, args.warmup,
                                              native)
    gsec, _, _, _, gout, gtop = run_steps(torch, handle, ix, d_reads, d_off, n,
                                          native.MH_CLASSIFY_BATCHED | native.MH_CLASSIFY_GRAPH,
                                          args.steps, args.warmup, native)
    r["reads_per_s"] = round(n / min(sec, gsec), 1)
    r["ms_per_step"] = round(min(sec, gsec) * 1e3, 4)
    r["graph_replay_same_results"] = same_results(native, gout, gtop, out, top)
    r["probes_per_s"] = round(ctr["probes"] / min(sec, gsec), 1)
    r["direct_hit_strands"] = ctr["direct"]
    log(f"workload 2 batched: {min(sec, gsec) * 1e3:.3f} ms/step, {r['reads_per_s'] / 1e6:.1f} M "
        "reads/s")
    rows = kernel_table(tim, kctr, tiles, tiers, 150)
    dom = max(rows, key=lambda x: x["ms"])
    traffic, src = profiled_traffic("k_classify_lane16_workload2_traffic.json", dom["kernel"])
    r["roofline"] = roofline_of(dom, traffic, src)
    if not args.no_plain:
        psec, ptim, _, _, pout, ptop = run_steps(torch, handle, ix, d_reads, d_off, n, 0,
                                                 args.steps, args.warmup, native)
        r["plain"] = {"reads_per_s": round(n / psec, 1), "ms_per_step": round(psec * 1e3, 4),
                      "same_results": same_results(native, pout, ptop, out, top)}
    pre_sec, ptim, _, _, pout, ptop = run_steps(
        torch, handle, ix, d_reads, d_off, n,
        native.MH_CLASSIFY_BATCHED | native.MH_CLASSIFY_PREPROBE, args.steps, args.warmup, native)
    probe_ms = float(np.mean([t_["probe_ms"] for t_ in ptim]))
    palg = 24 * ctr["probes"] + 2 * int(qo[-1]) + 28 * ctr["queries"]
    r["preprobe"] = {"reads_per_s": round(n / pre_sec, 1), "probe_kernel_ms": round(probe_ms, 4),
                     "same_results": same_results(native, pout, ptop, out, top),
                     "probe_alg_bytes": palg,
                     "probe_alg_frac_of_peak": round(palg / (probe_ms / 1e3) / 1e9 / HBM_PEAK_GBS,
                                                     4)}
    ptraf, _ = profiled_traffic("k_probe_workload2_traffic.json", "k_probe")
    if ptraf:
        r["preprobe"]["probe_traffic_frac_of_peak"] = round(
            ptraf / (probe_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4)
    if not args.no_fasta:
        r["fasta_ingest"] = fasta_leg(torch, native, handle, qb, qo, log)
    if not args.no_mhm:
        r["mhm_load"] = mhm_leg(torch, native, handle, ix, info, rb, ro, w.threshold, d_reads,
                                d_off, n, out, top, log)
    del d_reads, d_off
    mixed = None
    if not args.no_mixed:
        wm = D.workload(3)
        mqb, mqo, _ = D.make_reads(wm.reads, rb, ro, wm.seed)
        md_reads, md_off = device_tensors(torch, mqb, mqo)
        nm = len(mqo) - 1
        bs, btim, _, bk, bo, bt = run_steps(torch, handle, ix, md_reads, md_off, nm,
                                            native.MH_CLASSIFY_BATCHED, max(2, args.steps // 2),
                                            1, native)
        ps, _, _, _, po, pt = run_steps(torch, handle, ix, md_reads, md_off, nm, 0,
                                        max(2, args.steps // 2), 1, native)
        # bases per strand: the mean length of each length tile's reads (the
        # batch mean for the overflow tiers)
        lens = np.diff(mqo.astype(np.int64))
        mean = int(mqo[-1]) // nm
        per_row, lo = [], 0
        for i in range(8):
            if i < len(tiles):
                sel = lens[(lens > lo) & (lens <= tiles[i][0])]
                per_row.append(int(sel.mean()) if len(sel) else mean)
                lo = tiles[i][0]
            else:
                per_row.append(mean)
        per_row += [mean] * 4
        mrows = kernel_table(btim, bk, tiles, tiers, per_row)
        mdom = max(mrows, key=lambda x: x["ms"])
        mtraffic, msrc = profiled_traffic("k_classify_mixed_lengths_traffic.json", mdom["kernel"])
        mixed = {"reads": nm, "bases": int(mqo[-1]), "batched_reads_per_s": round(nm / bs, 1),
                 "plain_reads_per_s": round(nm / ps, 1), "speedup": round(ps / bs, 3),
                 "same_results": same_results(native, bo, bt, po, pt),
                 "roofline": roofline_of(mdom, mtraffic, msrc),
                 "roofline_note": "read bases per strand: the mean length of the tile's reads"}
        log(f"mixed-length: batched {nm / bs / 1e6:.3f} M reads/s, plain {nm / ps / 1e6:.3f} "
            f"M reads/s, x{ps / bs:.2f}")
        del md_reads, md_off
    if not args.no_cpu:
        from oracle import oracle as O
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        t = time.perf_counter()
        oix = O.Index(rb, ro, w.threshold)
        build_s = time.perf_counter() - t
        r["cpu_baseline"] = cpu_baseline(oix, qb, qo, args.w2_cpu_sample, threads, out, top,
                                         args.w2_cpu_sample // 10, build_s, "full index")
        log(f"workload 2 cpu: {r['cpu_baseline']['value'] / 1e3:.1f} k reads/s, bit-exact "
            f"{r['cpu_baseline']['bit_exact_vs_gpu']}")
        del oix
    ix.free()
    handle.close()
    return r, mixed


def workload4_leg(args, torch, native, D):
    """Served request stream of workload 4: 20,000 references ~2 Gbp, k = 16;
    requests of floor(10^U(0,4)) reads, 100-5,000 bp, 2 % substitutions +
    0.5 % indels; served and plain on the same stream."""
    w = D.workload(4, 1.0)
    rb, ro = D.make_refs(w.refs, w.seed)
    pool = D.ReadSpec(n=300_000, lengths=D.LengthSpec(mode=2, a=100, b=5000), sub_rate=0.02,
                      indel_rate=0.005)
    qb, qo, _ = D.make_reads(pool, rb, ro, w.seed)
    h = native.Handle(0, 8192)
    ix = h.build_index(rb, ro, w.threshold)
    rng = np.random.default_rng(w.seed)
    reqs, sizes = make_requests(rng, qb, qo, args.w4_requests)
    log(f"workload 4: {len(ro) - 1} refs, {len(rb) / 1e9:.2f} Gbp; {len(reqs)} requests")
    sv = serve_stream(native, ix, reqs, sizes, (0.25, 0.5, 0.9), rng, log)
    sv["workload"] = (f"{len(ro) - 1} refs {len(rb) / 1e9:.2f} Gbp k=16; requests of "
                      f"floor(10^U(0,4)) reads (mean {sizes.mean():.0f}), 100-5000 bp, "
                      "2% subs + 0.5% indels, Poisson arrivals")
    if not args.no_plain:
        sv["plain_reads_per_s"] = plain_stream(ix, reqs[: args.plain_requests],
                                               sizes[: args.plain_requests])
        sv["served_over_plain"] = round(sv["saturation"]["reads_per_s"] /
                                        sv["plain_reads_per_s"], 2)
        log(f"workload 4 plain on the request stream: {sv['plain_reads_per_s'] / 1e6:.2f} M "
            "reads/s")
    ix.free()
    h.close()
    return sv


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--w5-scale", type=float, default=1.0, help="fraction of workload 5's index")
    ap.add_argument("--w5-reads", type=int, default=10_000_000)
    ap.add_argument("--w5-form", choices=("auto", "batched", "preprobe"), default="auto",
                    help="workload 5's batched call: the faster of the two forms (auto), "
                         "or only the one named (for profiler runs)")
    ap.add_argument("--w5-cpu-sample", type=int, default=20_000)
    ap.add_argument("--w2-cpu-sample", type=int, default=1_000_000)
    ap.add_argument("--max-read-len", type=int, default=4096)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--serve-requests", type=int, default=4000,
                    help="requests of the headline's served stream")
    ap.add_argument("--serve-batch", type=int, default=65536,
                    help="most reads per served batch (headline stream)")
    ap.add_argument("--serve-slots", type=int, default=3,
                    help="served batches on the device at once (headline stream)")
    ap.add_argument("--w4-requests", type=int, default=300)
    ap.add_argument("--plain-requests", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-w2", action="store_true", help="skip workload 2 and the mixed batch")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-w4", action="store_true")
    ap.add_argument("--no-mhm", action="store_true")
    ap.add_argument("--no-fasta", action="store_true")
    ap.add_argument("--only-w2", action="store_true",
                    help="profiling run: workload 2 (and the mixed batch unless --no-mixed) "
                         "alone; the line's value is workload 2's")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also copy 1 GiB device to device once, the reference kernel for "
                         "scaling HBM byte counters in a profiler run")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs the classification path on a "
              "single device and has no multi-GPU mode (--gpus must be 1).", file=sys.stderr)
        sys.exit(2)

    import torch
    from mirage_amd import datagen as D
    from mirage_amd import native

    if args.calibrate_copy:
        a = torch.ones(1 << 30, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        log("copy calibration: 1 GiB read + 1 GiB written")
        del a, b
    if args.only_w2:
        w2, mixed = workload2_leg(args, torch, native, D)
        print(json.dumps({
            "metric": "reads classified/sec on 1 MI355X (profiling run, workload 2 only)",
            "value": w2["reads_per_s"], "unit": "reads/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": w2["ms_per_step"], "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "u8", "data": "synthetic (seeded)",
            "config": {"workload": w2["workload"]}, "workload2": w2, "mixed_length": mixed}),
            flush=True)
        return
    h5 = headline_leg(args, torch, native, D)
    w2 = mixed = w4 = None
    if not args.no_w2:
        w2, mixed = workload2_leg(args, torch, native, D)
    if not args.no_w4 and not args.no_serve:
        w4 = workload4_leg(args, torch, native, D)

    n = h5["n"]
    line = {
        "metric": "reads classified/sec on 1 MI355X, batched+served vs plain; p50/p99 latency",
        "value": round(n / h5["sec"], 1),
        "unit": "reads/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(h5["sec"] * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic (seeded)",
        "config": {"workload": f"large repetitive index: {h5['refs']} refs, "
                               f"{h5['ref_bases'] / 1e9:.2f} Gbp (30% holding 1-3 of 1,000 shared "
                               f"1-20 kb segments), k=20 (T=38); {n} reads x 150 bp, 1% subs; "
                               "batched + served on one GPU",
                   "mode": "batched (length-bucketed persistent kernels), best of eager and "
                           "hipGraph replay of the faster of the call with and without the "
                           "probe pass",
                   "form": h5["form"],
                   "reads": n, "index_entries": h5["index_entries"],
                   "index_GiB": h5["index_GiB"], "index_build_s": h5["index_build_s"]},
        "probes_per_s": round(h5["counters"]["probes"] / h5["sec"], 1),
        "entries_per_probe": h5["entries_per_probe"],
        "escalations": h5["escalations"],
        "graph_replay_same_results": h5["graph_replay_same_results"],
        "forms_same_results": h5["forms_same_results"],
        "eager_forms_ms_per_step": h5["eager_forms_ms_per_step"],
        "call_ms": h5["call_ms"],
        "eager_ms_per_step": h5["eager_ms_per_step"],
        "graph_ms_per_step": h5["graph_ms_per_step"],
        "tiles": h5["tiles"],
        "tiers": h5["tiers"],
        "kernels": h5["kernels"],
        "plain": h5.get("plain"),
        "served": h5.get("served"),
        "roofline": h5["roofline"],
        "cpu_baseline": h5.get("cpu_baseline"),
        "workload2": w2,
        "mixed_length": mixed,
        "served_workload4": w4,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
