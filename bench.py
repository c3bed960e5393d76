#!/usr/bin/env python3
"""Benchmark of the read-classification path on one MI355X.

Workload: 5,000 references (log-normal lengths, median 50 kb, sigma 1.0, capped
at 2 Mb, GC 40-60 %), k = 16 (T = 30), 1,000,000 reads x 150 bp (85 % sampled
with 2 % substitutions, 5 % exact substrings, 10 % random), seeded synthetic
data.  A step classifies the whole batch (both strands of every read: probe,
filters, chains, scoring, best hit, strand merge) with the reads already in
HBM.  Prints one JSON line; progress goes to stderr.

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


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def device_tensors(torch, qb, qo):
    d_reads = torch.from_numpy(qb).to("cuda")
    d_off = torch.from_numpy(qo.view(np.int64)).to("cuda")
    return d_reads, d_off


def run_steps(torch, handle, ix, d_reads, d_off, n, flags, steps, warmup, native):
    """Warm-up, then `steps` timed classify calls bracketed by synchronisation.
    Returns (seconds per step, per-step kernel timings, counters)."""
    d_out = torch.empty(n * native.READ_RESULT_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
    top_cap = 16 * n + 1024
    d_top = torch.empty(top_cap * 4, dtype=torch.uint8, device="cuda")
    stream = torch.cuda.ExternalStream(handle.stream_ptr())
    for _ in range(warmup):
        tn, _ = ix.classify_device(d_reads, d_off, n, d_out, d_top, top_cap, flags=flags)
        assert tn <= top_cap
    torch.cuda.synchronize()
    timings = []
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
    out = d_out.cpu().numpy().view(native.READ_RESULT_DTYPE)
    top = d_top.cpu().numpy().view(np.uint32)
    return max(wall, ev_s) / steps, timings, counters, out, top


def serving_leg(native, D, scale, n_requests, loads, log):
    """Served request stream (workload 4): 20,000 references ~2 Gbp (scaled),
    k = 16; requests of floor(10^U(0,4)) reads, 100-5,000 bp, 2 % substitutions
    + 0.5 % indels, Poisson arrivals.  Saturation throughput first, then the
    given fractions of it as offered loads; latency is submit -> results
    ready, measured by the server."""
    import threading
    w = D.workload(4, scale)
    t = time.perf_counter()
    rb, ro = D.make_refs(w.refs, w.seed)
    pool = D.ReadSpec(n=300_000, lengths=D.LengthSpec(mode=2, a=100, b=5000), sub_rate=0.02,
                      indel_rate=0.005)
    qb, qo, _ = D.make_reads(pool, rb, ro, w.seed)
    log(f"serving: {len(ro) - 1} refs, {len(rb) / 1e9:.2f} Gbp, read pool {len(qo) - 1} "
        f"({time.perf_counter() - t:.1f} s)")
    h = native.Handle(0, 8192)
    t = time.perf_counter()
    ix = h.build_index(rb, ro, w.threshold)
    info = ix.info()
    log(f"serving index: {info['n_entries']} entries, {info['device_bytes'] / 2**30:.1f} GiB, "
        f"{time.perf_counter() - t:.1f} s")
    rng = np.random.default_rng(w.seed)
    sizes = np.floor(10 ** rng.uniform(0, 4, n_requests)).astype(int)
    starts = rng.integers(0, len(qo) - 1 - sizes)
    reqs = [(qb[int(qo[a]):int(qo[a + n])], qo[a:a + n + 1] - qo[a]) for a, n in zip(starts, sizes)]
    mean_reads = float(sizes.mean())

    def run(rate):
        srv = native.Server(ix, hbm_budget=16 << 30, max_inflight=3, max_batch_reads=65536,
                            flags=0)
        ids, lat = [], []
        waiter_done = threading.Event()

        def waiter():
            for i in range(len(reqs)):
                while i >= len(ids):
                    time.sleep(0.0002)
                _, _, inf = srv.wait(ids[i])
                lat.append((inf["done_s"] - inf["submit_s"], inf["submit_s"], inf["done_s"]))
            waiter_done.set()

        th = threading.Thread(target=waiter)
        th.start()
        t0 = time.perf_counter()
        gaps = rng.exponential(1.0 / rate, len(reqs)) if rate else np.zeros(len(reqs))
        nxt = t0
        for (b, o), g in zip(reqs, gaps):
            nxt += g
            d = nxt - time.perf_counter()
            if d > 0:
                time.sleep(d)
            ids.append(srv.submit(b, o))
        th.join()
        srv.stop()
        l = np.array([x[0] for x in lat])
        span = max(x[2] for x in lat) - min(x[1] for x in lat)
        return {"reads_per_s": float(sizes.sum() / span), "p50_ms": float(np.percentile(l, 50) * 1e3),
                "p99_ms": float(np.percentile(l, 99) * 1e3), "requests": len(reqs)}

    sat = run(0)
    log(f"serving saturation: {sat['reads_per_s'] / 1e6:.2f} M reads/s")
    out = {"workload": f"{len(ro) - 1} refs {len(rb) / 1e9:.2f} Gbp k=16; requests of "
                       f"floor(10^U(0,4)) reads (mean {mean_reads:.0f}), 100-5000 bp, "
                       "2% subs + 0.5% indels, Poisson arrivals",
           "saturation": sat, "loads": []}
    for f in loads:
        rate = f * sat["reads_per_s"] / mean_reads
        r = run(rate)
        r["offered_fraction"] = f
        r["offered_requests_per_s"] = rate
        out["loads"].append(r)
        log(f"serving load {f:.2f}: {r['reads_per_s'] / 1e6:.2f} M reads/s, p50 "
            f"{r['p50_ms']:.2f} ms, p99 {r['p99_ms']:.2f} ms")
    ix.free()
    h.close()
    return out


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
    log(f"mhm: wrote {size / 2**30:.2f} GiB in {save_s:.2f} s ({finfo['n_buckets']} buckets, "
        f"{finfo['record_bytes'] / 2**30:.2f} GiB of records)")
    loads, same_index, same_results = [], True, None
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
                same_results = bool(
                    np.array_equal(o2["highest"].view(np.uint32), out["highest"].view(np.uint32))
                    and native.tie_sets(o2, t2) == native.tie_sets(out, top))
                del lout, ltop
            lx.free()
    finally:
        os.remove(path)
    best = min(loads)
    log(f"mhm: load {best:.3f} s ({size / best / 1e9:.2f} GB/s of file), index same "
        f"{same_index}, results same {same_results}")
    return {"file_bytes": size, "toc_entries": finfo["n_buckets"],
            "record_bytes": finfo["record_bytes"], "save_s": round(save_s, 3),
            "load_s": round(best, 4),
            "load_s_all": [round(x, 4) for x in loads],
            "load_file_GBps": round(size / best / 1e9, 3), "dir": d,
            "index_same_as_build": bool(same_index), "results_same": same_results}


def fasta_leg(torch, native, handle, qb, qo, out, top, log):
    """The reads as 61-column FASTA text in HBM, parsed on the device
    (mh_fasta_parse) five times: time per parse and GB/s of text; the parsed
    reads must equal the generated ones."""
    n = len(qo) - 1
    lens = np.diff(qo)
    if not np.all(lens == lens[0]):
        return {"skipped": "reads of one length only"}
    m = int(lens[0])
    seq = qb.reshape(n, m)
    cuts = list(range(61, m, 61))
    parts = np.split(seq, cuts, axis=1)
    nl = np.full((n, 1), ord("\n"), np.uint8)
    body = np.concatenate([x for p in parts for x in (p, nl)], axis=1)
    heads = [b">read%d\n" % i for i in range(n)]
    text = b"".join(h + body[i].tobytes() for i, h in enumerate(heads))
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


def workload5_leg(torch, native, D, log, scale=1.0, batch=1_000_000):
    """Workload 5 at full size (opt-in): 200,000 references of ~8 Gbp with
    inserted repeat segments, k = 20 (T = 38), 10M x 150 bp reads.  The
    index is built on the device; tiles are tuned on the first batch; the
    tuned and default tiles must give the same results on it; then every
    batch of `batch` reads is classified (batched form), timed end to end."""
    w = D.workload(5, scale)
    t = time.perf_counter()
    rb, ro, qb, qo, _ = D.generate(w)
    gen_s = time.perf_counter() - t
    n = len(qo) - 1
    log(f"workload 5: {len(rb) / 1e9:.2f} Gbp in {len(ro) - 1} refs, {n} reads "
        f"(generated in {gen_s:.1f} s)")
    h = native.Handle(0, 32768)
    t = time.perf_counter()
    ix = h.build_index(rb, ro, w.threshold)
    build_s = time.perf_counter() - t
    info = ix.info()
    log(f"workload 5 index: {info['n_entries']:,} entries, {info['device_bytes'] / 2**30:.1f} GiB "
        f"in HBM, built in {build_s:.1f} s (host copy included)")
    del rb
    d_reads, d_off = device_tensors(torch, qb, qo)
    nb = (n + batch - 1) // batch
    top_cap = 64 * batch + 1024
    d_out = torch.empty(batch * native.READ_RESULT_DTYPE.itemsize, dtype=torch.uint8,
                        device="cuda")
    d_top = torch.empty(top_cap * 4, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()

    def span(b):
        lo = b * batch
        return d_off[lo:min(lo + batch, n) + 1], min(batch, n - lo)

    def classify(off_b, n_b):
        # MH_ERR_OVERFLOW: the library grew its tie pool (or the output pool
        # is short); the call is made again.
        for attempt in range(4):
            try:
                return ix.classify_device(d_reads, off_b, n_b, d_out, d_top, top_cap,
                                          flags=native.MH_CLASSIFY_BATCHED)
            except native.MirageError as e:
                if e.status != native.MH_ERR_OVERFLOW or attempt == 3:
                    raise

    def run0():
        off0, n0 = span(0)
        tn, _ = classify(off0, n0)
        assert tn <= top_cap
        out = d_out.cpu().numpy().view(native.READ_RESULT_DTYPE)[:n0].copy()
        top = d_top.cpu().numpy().view(np.uint32).copy()
        return out, native.tie_sets(out, top), h.counters()

    default_tiles = h.get_tiles()
    out_d, ties_d, ctr = run0()
    t = time.perf_counter()
    off0, n0 = span(0)
    tiles = h.autotune(ix, d_reads, off0, n0, iters=1)
    tune_s = time.perf_counter() - t
    out_t, ties_t, _ = run0()
    same = bool(np.array_equal(out_d["highest"].view(np.uint32), out_t["highest"].view(np.uint32))
                and np.array_equal(out_d["strand"], out_t["strand"]) and ties_d == ties_t)
    log(f"workload 5 tiles tuned in {tune_s:.1f} s: {tiles[:2]} (default {default_tiles[:2]}); "
        f"results same {same}")
    needed = 0
    for b in range(nb):  # checked pass: the top-tie pool holds every batch
        off_b, n_b = span(b)
        tn, _ = classify(off_b, n_b)
        needed = max(needed, tn)
    assert needed <= top_cap
    torch.cuda.synchronize()
    t = time.perf_counter()
    for b in range(nb):
        off_b, n_b = span(b)
        ix.classify_device(d_reads, off_b, n_b, d_out, d_top, top_cap,
                           flags=native.MH_CLASSIFY_BATCHED, wait=False, sync_inputs=False)
    torch.cuda.synchronize()
    sec = time.perf_counter() - t
    probes, entries = ctr["probes"], ctr["entries"]
    log(f"workload 5: {n / sec / 1e6:.2f} M reads/s over {nb} batches of {batch} "
        f"({sec * 1e3:.1f} ms), {entries / max(probes, 1):.1f} entries per probe, "
        f"{ctr['overflow']} of {2 * n0} strands of batch 0 in the overflow tiers")
    res = {"scale": scale, "reads": n, "refs": len(ro) - 1, "ref_bases": int(ro[-1]), "threshold": w.threshold,
           "index_entries": info["n_entries"], "index_GiB": round(info["device_bytes"] / 2**30, 2),
           "index_build_s": round(build_s, 2), "batch_reads": batch, "batches": nb,
           "reads_per_s": round(n / sec, 1), "ms_total": round(sec * 1e3, 2),
           "entries_per_probe": round(entries / max(probes, 1), 2),
           "overflow_strands_batch0": ctr["overflow"], "tiles": [list(x) for x in tiles],
           "default_tiles_same_results": same, "tune_s": round(tune_s, 1),
           "top_pool_needed_max": int(needed)}
    ix.free()
    h.close()
    return res


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(rb, ro, qb, qo, threshold, sample, threads, gpu_out, gpu_top, sample1):
    """The oracle (C++ restatement of the reference CPU path) on a bounded
    sample of the same reads: one warm-up, then the median of 5 runs with
    `threads` threads, and the same on a smaller sample with 1 thread.  Also
    checks the GPU results on the sample."""
    from oracle import oracle as O
    from mirage_amd import native
    t = time.perf_counter()
    oix = O.Index(rb, ro, threshold)
    build_s = time.perf_counter() - t

    def timed(n, nt):
        sq, so = qb[: int(qo[n])], qo[: n + 1]
        oix.classify(sq, so, n_threads=nt)  # warm-up
        runs, res = [], None
        for _ in range(5):
            t0 = time.perf_counter()
            res = oix.classify(sq, so, n_threads=nt)
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
    return {"reads_per_s": n / med, "seconds": med, "reads": n, "index_build_s": build_s,
            "bit_exact": bool(same), "reads_per_s_1thread": n1 / med1, "reads_1thread": n1}


def profiled_traffic(fname, kernel, scale):
    """(HBM bytes per launch, source) from a committed profiler summary of
    `kernel` on the full-size workload, or None."""
    path = os.path.join(ROOT, "profiles", fname)
    if scale != 1.0 or not os.path.exists(path):
        return None
    with open(path) as f:
        tj = json.load(f)
    if tj.get("kernel") != kernel:
        return None
    return int(tj["traffic_bytes"]), tj.get("source")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--scale", type=float, default=1.0, help="fraction of the 1M reads")
    ap.add_argument("--cpu-sample", type=int, default=1_000_000)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-preprobe", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--mixed-reads", type=int, default=2_000_000)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-mhm", action="store_true")
    ap.add_argument("--no-fasta", action="store_true")
    ap.add_argument("--workload5", action="store_true",
                    help="also run workload 5 at full size (200,000 refs, ~8 Gbp, 10M reads; "
                         "about 80 GB of HBM, a few minutes)")
    ap.add_argument("--workload5-scale", type=float, default=1.0)
    ap.add_argument("--serve-scale", type=float, default=1.0)
    ap.add_argument("--serve-requests", type=int, default=300)
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

    w = D.workload(2, args.scale)
    t = time.perf_counter()
    rb, ro, qb, qo, truth = D.generate(w)
    n = len(qo) - 1
    log(f"data: {len(ro) - 1} refs, {len(rb) / 1e6:.1f} Mbp, {n} reads "
        f"({time.perf_counter() - t:.1f} s)")
    handle = native.Handle(0, 32768)
    t = time.perf_counter()
    ix = handle.build_index(rb, ro, w.threshold)
    torch.cuda.synchronize()
    index_s = time.perf_counter() - t
    info = ix.info()
    log(f"device index: {info['n_entries']} entries, {info['device_bytes'] / 2**30:.2f} GiB, "
        f"{index_s:.2f} s")
    d_reads, d_off = device_tensors(torch, qb, qo)

    log("batched: warm-up + timed steps")
    sec, timings, ctr, out, top = run_steps(torch, handle, ix, d_reads, d_off, n,
                                            native.MH_CLASSIFY_BATCHED, args.steps, args.warmup,
                                            native)
    log(f"batched (eager launches): {sec * 1e3:.3f} ms/step, {n / sec / 1e6:.2f} M reads/s")
    gsec, _, _, gout, gtop = run_steps(torch, handle, ix, d_reads, d_off, n,
                                       native.MH_CLASSIFY_BATCHED | native.MH_CLASSIFY_GRAPH,
                                       args.steps, args.warmup, native)
    graph_same = (np.array_equal(gout["highest"].view(np.uint32), out["highest"].view(np.uint32))
                  and native.tie_sets(gout, gtop) == native.tie_sets(out, top))
    log(f"batched (hipGraph replay): {gsec * 1e3:.3f} ms/step, {n / gsec / 1e6:.2f} M reads/s, "
        f"same results {graph_same}")
    batched_rps = n / min(sec, gsec)
    sec = min(sec, gsec)

    # Optional form: a separate probe pass hashes every window of the 150 bp
    # strands and looks up its bucket before the tile kernel runs.
    pre = None
    if not args.no_preprobe:
        psec_, ptim_, _, pout_, ptop_ = run_steps(
            torch, handle, ix, d_reads, d_off, n,
            native.MH_CLASSIFY_BATCHED | native.MH_CLASSIFY_PREPROBE, args.steps, args.warmup,
            native)
        same = (np.array_equal(pout_["highest"].view(np.uint32), out["highest"].view(np.uint32))
                and native.tie_sets(pout_, ptop_) == native.tie_sets(out, top))
        probe_ms = float(np.mean([t_["probe_ms"] for t_ in ptim_]))
        # k_probe per launch: the directory pair (16 B) and the record (8 B) of
        # every probe, the read bytes of both strands, and per strand its list
        # entry and offsets (20 B) and status word (8 B).
        palg = 24 * ctr["probes"] + 2 * int(qo[-1]) + 28 * ctr["queries"]
        pre = {"reads_per_s": n / psec_, "ms_per_step": psec_ * 1e3, "same_results": bool(same),
               "probe_kernel_ms": round(probe_ms, 4),
               "tile_kernel_ms": round(float(np.mean([t_["bucket_ms"][0] for t_ in ptim_])), 4),
               "probe_alg_bytes": palg,
               "probe_achieved_GBs": round(palg / (probe_ms / 1e3) / 1e9, 1)}
        # measured HBM bytes of the probe pass (profiler counters, committed
        # summary) over the pass time measured here
        ptraf = profiled_traffic("k_probe_workload2_traffic.json", "k_probe", args.scale)
        if ptraf:
            pre["probe_traffic_bytes"] = ptraf[0]
            pre["probe_traffic_GBs"] = round(ptraf[0] / (probe_ms / 1e3) / 1e9, 1)
            pre["probe_traffic_frac_of_peak"] = round(pre["probe_traffic_GBs"] / HBM_PEAK_GBS, 4)
        log(f"pre-probe form: {psec_ * 1e3:.3f} ms/step, probe pass {probe_ms:.3f} ms "
            f"({pre['probe_achieved_GBs']} GB/s algorithmic), same results {same}")
        if psec_ < sec:
            batched_rps, sec = n / psec_, psec_
    if args.calibrate_copy:
        a = torch.ones(1 << 30, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        log("copy calibration: 1 GiB read + 1 GiB written")
        del a, b

    plain = None
    if not args.no_plain:
        psec, ptim, _, pout, ptop = run_steps(torch, handle, ix, d_reads, d_off, n, 0, args.steps,
                                              args.warmup, native)
        same = (np.array_equal(pout["highest"].view(np.uint32), out["highest"].view(np.uint32))
                and native.tie_sets(pout, ptop) == native.tie_sets(out, top))
        plain = {"reads_per_s": n / psec, "ms_per_step": psec * 1e3, "same_results": bool(same),
                 "kernel_ms": float(np.mean([t_["bucket_ms"][0] for t_ in ptim]))}
        log(f"plain: {psec * 1e3:.3f} ms/step, {n / psec / 1e6:.2f} M reads/s")

    mixed = None
    if not args.no_mixed:
        wm = D.workload(3)
        wm.reads.n = args.mixed_reads
        mqb, mqo, _ = D.make_reads(wm.reads, rb, ro, wm.seed)
        md_reads, md_off = device_tensors(torch, mqb, mqo)
        nm = len(mqo) - 1
        bs, _, _, bo, bt = run_steps(torch, handle, ix, md_reads, md_off, nm,
                                     native.MH_CLASSIFY_BATCHED, max(2, args.steps // 2), 1, native)
        ps, _, _, po, pt = run_steps(torch, handle, ix, md_reads, md_off, nm, 0,
                                     max(2, args.steps // 2), 1, native)
        same = (np.array_equal(bo["highest"].view(np.uint32), po["highest"].view(np.uint32))
                and native.tie_sets(bo, bt) == native.tie_sets(po, pt))
        mixed = {"reads": nm, "bases": int(mqo[-1]), "batched_reads_per_s": nm / bs,
                 "plain_reads_per_s": nm / ps, "speedup": ps / bs, "same_results": bool(same)}
        log(f"mixed-length: batched {nm / bs / 1e6:.3f} M reads/s, plain {nm / ps / 1e6:.3f} "
            f"M reads/s, x{ps / bs:.2f}")
        del md_reads, md_off

    wl5 = None
    if args.workload5:
        wl5 = workload5_leg(torch, native, D, log, args.workload5_scale)

    fasta = None
    if not args.no_fasta:
        fasta = fasta_leg(torch, native, handle, qb, qo, out, top, log)

    mhm = None
    if not args.no_mhm:
        mhm = mhm_leg(torch, native, handle, ix, info, rb, ro, w.threshold, d_reads, d_off, n,
                      out, top, log)

    # Roofline of the dominant kernel: the tile kernel of the shortest length
    # bucket (16-lane teams, reads up to 200 bp), which holds every read of
    # this workload.  Algorithmic bytes per launch: a 16-byte directory pair
    # per probe, 8 bytes per bucket entry read, 2 bytes per extension
    # comparison, the read bases of both strands, and a 32-byte result per
    # strand.
    kernel_ms = float(np.mean([t_["bucket_ms"][0] for t_ in timings]))
    alg_bytes = (16 * ctr["probes"] + 8 * ctr["entries"] + 2 * ctr["ext_compares"] +
                 2 * int(qo[-1]) + 32 * ctr["queries"])
    achieved = alg_bytes / (kernel_ms / 1e3) / 1e9
    kernel_name = "k_classify<LaneTeam<16>> (reads <= 200 bp)"
    # HBM bytes per launch of this kernel from the profiler's FETCH_SIZE and
    # WRITE_SIZE counters, scaled by what they read for a 1 GiB copy (the
    # summary in profiles/ says how it was measured).
    traffic, traffic_src = profiled_traffic("k_classify_lane16_workload2_traffic.json",
                                            kernel_name, args.scale) or (None, None)
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                "traffic_source": traffic_src, "kernel": kernel_name,
                "kernel_ms": round(kernel_ms, 4), "alg_bytes_per_launch": alg_bytes}

    serve = None
    if not args.no_serve:
        serve = serving_leg(native, D, args.serve_scale, args.serve_requests, (0.25, 0.5, 0.9),
                            log)

    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        log(f"cpu baseline: oracle on {min(args.cpu_sample, n)} reads, {threads} threads")
        cb = cpu_baseline(rb, ro, qb, qo, w.threshold, args.cpu_sample, threads, out, top,
                          args.cpu_sample // 10)
        cpu = {"value": round(cb["reads_per_s"], 1), "unit": "reads/s", "cores": threads,
               "kind": "port",
               "sample": f"first {cb['reads']} reads of the same workload against the full "
                         f"index, median of 5 after 1 warm-up (oracle index build "
                         f"{cb['index_build_s']:.1f} s, not timed)",
               "bit_exact_vs_gpu": cb["bit_exact"],
               "value_1thread": round(cb["reads_per_s_1thread"], 1),
               "sample_1thread": f"first {cb['reads_1thread']} reads, median of 5",
               "host_nproc": os.cpu_count(), "host_cpu": cpu_model()}
        log(f"cpu: {cb['reads_per_s'] / 1e3:.1f} k reads/s, gpu results on the sample "
            f"bit-exact: {cb['bit_exact']}")

    line = {
        "metric": "reads classified/sec on 1 MI355X, batched+served vs plain; p50/p99 latency",
        "value": round(batched_rps, 1),
        "unit": "reads/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(sec * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic (seeded)",
        "config": {"workload": "5000 refs lognormal(50 kb, 1.0) GC 40-60%, k=16 (T=30), "
                               f"{n} reads x 150 bp, 2% subs, 5% exact, 10% random",
                   "mode": "batched (length-bucketed persistent kernels)",
                   "reads": n, "index_entries": info["n_entries"],
                   "index_build_s": round(index_s, 3)},
        "probes_per_s": round(ctr["probes"] / sec, 1),
        "direct_hit_strands": ctr["direct"],
        "graph_replay_same_results": bool(graph_same),
        "preprobe": pre,
        "plain": plain,
        "mixed_length": mixed,
        "mhm_load": mhm,
        "fasta_ingest": fasta,
        "workload5": wl5,
        "served": serve,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
