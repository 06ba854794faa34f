"""Benchmark: input Gbases/s of minimizer bucketing + per-sub-bucket k-mer
merge on one MI355X, inputs resident in HBM, outputs bit-exact with the
CPU oracle (checked on a sample before timing).

python bench.py [--gpus 1] [--steps K] [--warmup W] [--config 2] [--scale 1.0]

A step = one bucketing pass + one merge pass over the whole workload.
Prints one JSON line.
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

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def make_workload(config: int, scale: float):
    from ggcat_amd import synth

    if config == 1:
        return synth.config1(scale)
    if config == 2:
        return synth.config2(scale)
    if config == 4:
        return synth.config4(scale)
    raise SystemExit(f"unsupported config {config}")


def kernel_bytes(cnt: np.ndarray, key_words: int) -> dict:
    """Algorithmic HBM bytes per launch of each kernel.
    cnt: [input bytes, records, sum record bases, k-mer occurrences, big, solid,
    ..., stored records (counter 10)].
    The super-k-mer arena is sum ceil(l/4) + 8 B of metadata per record; the
    staging pass reads the input and writes every record once (into its
    unit's region); the compaction reads the regions and writes the stored
    records (identical records of a unit once) into the dense arena; the
    merge reads the stored arena and writes the solid entries."""
    in_b, n_rec, sum_len, _, _, n_solid = (float(x) for x in cnt[:6])
    n_stored = float(cnt[10]) if len(cnt) > 10 and cnt[10] else n_rec
    arena = (sum_len + 1.5 * n_rec) / 4.0 + 8.0 * n_rec
    stored = arena * (n_stored / n_rec) if n_rec else 0.0
    return {
        "bucket_stage": in_b + arena,
        "unit_compact": arena + stored,
        "merge": stored + n_solid * (8.0 * key_words + 8.0),
    }


# rocprofv3 FETCH_SIZE / WRITE_SIZE of one launch per kernel at full size,
# scaled by the counters' reading for a 1 GiB copy (profiles/ holds the
# commands and the per-kernel table)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_config2.json")
TRAFFIC_KERNELS = {"bucket_stage": ("gg::bucket_stage_kernel<16>",),
                   "unit_compact": ("gg::unit_dedup_kernel", "gg::unit_dedup_wide_kernel", "gg::unit_compact_kernel"),
                   "merge": ("gg::merge_lds_seq64_kernel",)}


def measured_traffic(workload: str, scale: float, kernel: str):
    """HBM bytes (GB, read + write) of one launch of `kernel` as measured by
    the committed counter runs of the same workload, else (None, None)."""
    if not workload.startswith("config2") or scale != 1.0 or not os.path.exists(TRAFFIC_FILE):
        return None, None
    table = json.load(open(TRAFFIC_FILE))
    parts = [table.get(name) for name in TRAFFIC_KERNELS.get(kernel, ())]
    if not parts or not all(parts):
        return None, None
    gb = sum(t["read_gb"] + t["write_gb"] for t in parts)
    return round(gb, 3), "profiles/hbm_traffic_config2.md (rocprofv3 --pmc, GB per launch)"


def kernel_rooflines(kb: dict, kern_ms: dict, workload: str, scale: float) -> dict:
    """Every timed kernel against the HBM peak: algorithmic bytes over its
    HIP-event time, and the counter-measured bytes over the same time."""
    out = {}
    for name, ms in kern_ms.items():
        alg = kb[name] / (ms * 1e-3) / 1e9
        traffic, _ = measured_traffic(workload, scale, name)
        out[name] = {"ms": round(float(ms), 3), "algorithmic_gb": round(kb[name] / 1e9, 3),
                     "achieved_gbs": round(alg, 1), "frac": round(alg / HBM_PEAK_GBS, 4),
                     "traffic_gb": traffic,
                     "traffic_gbs": round(traffic / (ms * 1e-3), 1) if traffic else None,
                     "traffic_frac": round(traffic / (ms * 1e-3) / HBM_PEAK_GBS, 4) if traffic else None}
    return out


def solid_matrix(keys, sub, mult, flags) -> np.ndarray:
    """Rows (sub, key_lo, key_hi, multiplicity, flags), sorted."""
    hi = keys[:, -1] if keys.shape[1] > 1 else np.zeros(keys.shape[0], np.uint64)
    m = np.stack([sub.astype(np.uint64), keys[:, 0].astype(np.uint64), hi.astype(np.uint64),
                  mult.astype(np.uint64), flags.astype(np.uint64)], 1)
    return m[np.lexsort(m.T[::-1])]


def parity_gate(ctx, wl, params, sample: np.ndarray, threads: int) -> dict:
    """GPU solid table of `sample` under the schedule that is about to be
    timed, against the oracle, bit for bit; exits on a mismatch."""
    from oracle import binding as O

    ctx.bucketing(sample, params)
    ctx.kmers_merge()
    split = int(ctx.counters()[4])
    s = ctx.solid()
    g = solid_matrix(s.keys, s.sub, s.mult, s.flags)
    t = time.perf_counter()
    o = O.bucketing(sample, wl.k, wl.m, wl.buckets_log, wl.second_buckets_log, threads=threads)
    o.merge(wl.k, wl.buckets_log, wl.second_buckets_log, wl.hash_kind, wl.min_mult, threads)
    os_ = o.solid()
    keys = np.stack([os_["key_lo"], os_["key_hi"]], 1) if g.shape[0] and s.keys.shape[1] > 1 else \
        os_["key_lo"].reshape(-1, 1)
    e = solid_matrix(keys, os_["sub"], os_["mult"], os_["flags"])
    dt = time.perf_counter() - t
    if not np.array_equal(g, e):
        raise SystemExit(f"[bench] parity gate FAILED: GPU solid table ({g.shape[0]} entries) differs from the "
                         f"oracle ({e.shape[0]} entries)")
    bases = int(np.count_nonzero(sample != ord("\n")))
    return {"bytes": int(sample.size), "solid_entries": int(g.shape[0]), "split_items": split,
            "oracle_s": round(dt, 2), "oracle_threads": threads, "oracle_gbases_s": round(bases / dt / 1e9, 4),
            "result": "bit-exact"}


def cpu_baseline(wl, threads: int) -> dict:
    """The oracle restatement on a bounded prefix of the same reads."""
    from oracle import binding as O

    target = 1_000_000_000 if wl.k <= 32 else 400_000_000  # bases in the sample (~10-20 s of CPU work)
    data = wl.data
    if data.size > target:
        cut = target
        while data[cut - 1] != ord("\n"):
            cut += 1
        data = data[:cut]
    bases = int(np.count_nonzero(data != ord("\n")))
    t0 = time.perf_counter()
    r = O.bucketing(data, wl.k, wl.m, wl.buckets_log, wl.second_buckets_log, threads=threads)
    r.merge(wl.k, wl.buckets_log, wl.second_buckets_log, wl.hash_kind, wl.min_mult, threads)
    dt = time.perf_counter() - t0
    return {"value": round(bases / dt / 1e9, 4), "unit": "Gbases/s", "cores": threads, "kind": "port",
            "sample": f"first {bases} bases ({data.size} bytes) of the same read set; oracle bucketing + merge, "
                      f"{threads} threads, {dt:.2f} s"}


def cpu_unitig_baseline(wl, threads: int) -> tuple[dict, dict]:
    """The oracle's unitig constructions on the same read-set shape at 1/40
    of the size (built with `threads` threads first): per-sub-bucket chains
    with `threads` workers over the sub-buckets, the joined maximal unitigs
    of the one global graph on a single thread."""
    from ggcat_amd import synth
    from oracle import binding as O

    sw = synth.config2(0.025)
    r = O.bucketing(sw.data, sw.k, sw.m, sw.buckets_log, sw.second_buckets_log, threads=threads)
    r.merge(sw.k, sw.buckets_log, sw.second_buckets_log, sw.hash_kind, sw.min_mult, threads)
    n = int(r.solid()["sub"].size)
    out = []
    for joined in (False, True):
        cores = 1 if joined else threads
        t = time.perf_counter()
        r.unitigs(sw.k, joined=joined, threads=cores)
        dt = time.perf_counter() - t
        out.append({"value": round(n / dt, 1), "unit": "solid k-mers/s", "cores": cores, "kind": "port",
                    "sample": f"{sw.name} ({sw.bases} bases, {n} solid k-mers); oracle "
                              f"{'joined' if joined else 'per-sub-bucket'} unitigs, {dt:.2f} s"})
    return out[0], out[1]


def end_to_end(ctx, wl, schedule: dict, target_bases: int = 1_000_000_000) -> dict:
    """File -> solid table: the first ~1 Gbase of the read set written as a
    FASTA file (one '>' header line per read), then ggcat_gpu_build_from_files
    timed as a whole: the file read into HBM through pinned buffers, parsed
    there, bucketing, merge.  The first call also allocates the context's
    buffers; the second (reported) reuses them.  The host reader path (parse
    on the CPU, then upload) is timed beside it and must give the same table."""
    import tempfile

    from ggcat_amd.native import BuildOptions

    data = wl.data
    cut = min(target_bases, data.size)
    while cut < data.size and data[cut - 1] != ord("\n"):
        cut += 1
    data = data[:cut]
    nl = data == ord("\n")
    n_reads = int(np.count_nonzero(nl))
    # record index of every byte; each record gains the two bytes ">\n" in front
    rec = np.zeros(data.size, np.uint32)
    np.cumsum(nl[:-1], dtype=np.uint32, out=rec[1:])
    fasta = np.empty(data.size + 2 * n_reads, np.uint8)
    pos = np.arange(data.size, dtype=np.uint64) + 2 * (rec.astype(np.uint64) + 1)
    fasta[pos] = data
    starts = np.concatenate([[0], np.flatnonzero(nl)[:-1] + 1]).astype(np.uint64)
    hdr = starts + 2 * np.arange(n_reads, dtype=np.uint64)
    fasta[hdr] = ord(">")
    fasta[hdr + 1] = ord("\n")
    del rec, pos
    fd, path = tempfile.mkstemp(suffix=".fa")
    os.close(fd)
    runs = []
    try:
        fasta.tofile(path)
        bases = int(data.size - n_reads)
        del fasta
        opts = BuildOptions(wl.k, wl.m, wl.min_mult, 0, 0, wl.hash_kind, 0)
        for host_reader in (0, 0, 1):
            ctx.set_tuning(**dict(schedule, host_reader=host_reader))
            t = time.perf_counter()
            used = ctx.build_from_files([path], opts)
            dt = time.perf_counter() - t
            c, tm = ctx.counters(), ctx.timings()
            runs.append({"wall_s": dt, "path": int(c[8]), "records": int(c[1]), "solid": int(c[5]),
                         "read_ms": float(tm[8]), "bucketing_ms": float(tm[5]), "merge_ms": float(tm[6])})
            log(f"[bench] end-to-end {'host reader' if host_reader else 'device parse'}: {dt:.3f} s "
                f"(input {tm[8]:.1f} ms, bucketing {tm[5]:.1f} ms, merge {tm[6]:.1f} ms)")
    finally:
        os.unlink(path)
        ctx.set_tuning(**schedule)
    first, warm, host = runs
    if warm["path"] != 1 or host["path"] != 2:
        raise RuntimeError(f"unexpected input paths {warm['path']} / {host['path']}")
    if (warm["records"], warm["solid"]) != (host["records"], host["solid"]):
        raise RuntimeError(f"device parse and host reader disagree: {warm} vs {host}")
    return {"gbases_s": round(bases / warm["wall_s"] / 1e9, 4), "wall_s": round(warm["wall_s"], 3),
            "first_call_wall_s": round(first["wall_s"], 3), "bases": bases,
            "file_bytes": int(data.size + 2 * n_reads), "input_ms": round(warm["read_ms"], 1),
            "device_bucketing_ms": round(warm["bucketing_ms"], 1), "device_merge_ms": round(warm["merge_ms"], 1),
            "host_reader_gbases_s": round(bases / host["wall_s"] / 1e9, 4),
            "host_reader_input_ms": round(host["read_ms"], 1), "solid": warm["solid"],
            "buckets_log": int(used.buckets_log),
            "sample": "first ~1 Gbase of the same read set as a FASTA file (in the page cache); file read into "
                      "HBM through pinned buffers + parse on the device + bucketing + merge, second call on the "
                      "context; host_reader_* = the same build with the host parser"}


def serve_measurement(ctx, rate_hz: float, n_requests: int) -> dict:
    """Colored query serving: the 256-genome colored graph (one color per
    genome, k=31 m=15 s=2) built and indexed in HBM, then a Poisson stream of
    requests of U{1..10000} queries with log-uniform lengths in 100..10000 bp,
    half sampled from the genomes with 2% substitutions, half i.i.d.; latency
    is submission to completion as seen by the polling loop."""
    from ggcat_amd import synth
    from ggcat_amd.native import Params
    from ggcat_amd.query import serve_poisson

    wl = synth.config3(1.0)
    starts, colors = wl.colors
    p = Params(wl.k, wl.m, wl.buckets_log, wl.second_buckets_log, 0, 0, wl.min_mult)
    t = time.perf_counter()
    ctx.bucketing(wl.data, p, colors=(starts, colors))
    ctx.kmers_merge()
    t_build = time.perf_counter() - t
    t = time.perf_counter()
    ctx.index_build()
    t_index = time.perf_counter() - t
    n_solid = int(ctx.counters()[5])
    gs = synth.config3_genomes(1.0)
    rng = np.random.default_rng(5)
    reqs = [synth.queries(gs, 500 + i, int(n), 100, 10_000, 0.02) for i, n in
            enumerate(rng.integers(1, 10_001, n_requests))]
    log(f"[bench] serving: graph of {len(gs)} genomes ({wl.bases} bases) built in {t_build:.2f} s, "
        f"{n_solid} solid k-mers indexed in {t_index * 1e3:.1f} ms; {n_requests} requests at {rate_hz}/s")
    ctx.serve_open(4, 2 << 30)
    results = {}
    try:
        rep = serve_poisson(ctx, reqs, rate_hz, results=results)
        # capacity: the same requests all at once
        burst = serve_poisson(ctx, reqs, 1e9)
    finally:
        ctx.serve_close()
    # each result equals the same request run alone (spot check of three)
    for i in (0, n_requests // 2, n_requests - 1):
        alone = ctx.query(reqs[i])
        if not all(np.array_equal(x, y) for x, y in zip(results[i], alone)):
            raise SystemExit("[bench] serving result differs from the request run alone")
    return {"graph": f"{wl.name}: {len(gs)} genomes, {wl.bases} bases, colored, k=31 m=15 s=2",
            "rate_hz": rate_hz, "requests": n_requests, "queries": rep.queries, "query_kmers": rep.kmers,
            "p50_ms": round(rep.percentile(50), 2), "p99_ms": round(rep.percentile(99), 2),
            "max_ms": round(max(rep.latencies_ms), 2), "wall_s": round(rep.wall_s, 2),
            "kmers_per_s": round(rep.kmers / rep.wall_s, 1), "max_admitted": rep.max_admitted,
            "burst_wall_s": round(burst.wall_s, 3), "burst_kmers_per_s": round(burst.kmers / burst.wall_s, 1),
            "burst_p99_ms": round(burst.percentile(99), 2),
            "graph_build_s": round(t_build, 3), "index_build_ms": round(t_index * 1e3, 1)}


def shape_sample(config: int, scale: float, full_bytes: int, cap: int) -> tuple[np.ndarray, str]:
    """A read set of the same configuration with its genome scaled so that it
    holds about `cap` bytes: coverage, read length, error rate and repeat
    content as at full size, so work items (and the items that need
    partitioned passes) look like the full-size ones."""
    f = min(1.0, cap / max(full_bytes, 1))
    s = make_workload(config, scale * f)
    return s.data, s.name


def host_input_steps(ctx, wl, params, steps: int, cnt_dev: np.ndarray) -> dict:
    """The step from the host array instead of HBM: ggcat_gpu_bucketing
    uploads the input in chunks on its copy stream while the staging pass
    runs over the tiles already resident.  Wall clock per step (upload
    included, the first call untimed); the result's counters must equal the
    device-resident run's."""
    ts = []
    for i in range(steps + 1):
        t = time.perf_counter()
        ctx.bucketing(wl.data, params)
        ctx.kmers_merge()
        ts.append(time.perf_counter() - t)
        log(f"[bench] host-input step {i}/{steps} done ({ts[-1] * 1e3:.1f} ms)")
    c = ctx.counters()
    # records, k-mer occurrences, solid entries (the stored-record count may
    # differ by a few: which records pass their unit's region, and so are
    # not deduplicated there, depends on the order the tiles run in)
    fields = (1, 3, 5)
    if any(int(c[i]) != int(cnt_dev[i]) for i in fields):
        raise SystemExit(f"[bench] host-input run differs from the device-resident run: "
                         f"{[int(c[i]) for i in fields]} vs {[int(cnt_dev[i]) for i in fields]}")
    tm = float(np.median(ts[1:]))
    return {"gbases_s": round(wl.bases / tm / 1e9, 4), "ms_per_step": round(tm * 1e3, 1),
            "staging_ms": round(float(ctx.timings()[0]), 1), "counters_equal": True,
            "note": "input in host memory; upload chunks of 64 MiB through pinned buffers, overlapped with the "
                    "staging pass; wall clock per step"}


def measure(ctx, wl, steps: int, warmup: int, tune: bool, check: bool, threads: int, scale: float,
            sample: np.ndarray | None = None, sample_name: str | None = None,
            schedule: dict | None = None, host_steps: int = 2) -> dict:
    """Upload the workload, pick the schedule on a sample (default: the first
    256 MiB of the reads), gate it against the oracle on that sample, then
    time `steps` bucketing + merge passes over the device-resident input.
    Leaves the last step's solid table in ctx."""
    from ggcat_amd.native import Params

    params = Params(wl.k, wl.m, wl.buckets_log, wl.second_buckets_log, 0, wl.hash_kind, wl.min_mult)
    t = time.perf_counter()
    dptr, nbytes = ctx.upload(wl.data)
    h2d_s = time.perf_counter() - t
    log(f"[bench] {wl.name}: input uploaded to HBM in {h2d_s * 1e3:.1f} ms")
    if sample is None:
        cut = min(wl.data.size, 256 << 20)
        while cut < wl.data.size and wl.data[cut - 1] != ord("\n"):
            cut += 1
        sample, sample_name = wl.data[:cut], f"first {cut} bytes of {wl.name}"
    cut = int(sample.size)
    tuning = None
    ctx.set_tuning()
    if schedule is not None:
        # a fixed schedule (e.g. the one an earlier run's tuner picked)
        from ggcat_amd.autotune import Schedule

        Schedule(**schedule).apply(ctx)
        tuning = {"schedule": schedule, "fixed": True}
    elif tune:
        # schedule picked on a bounded sample of the same input; every
        # candidate's solid table must have the same digest
        from ggcat_amd.autotune import autotune

        t = time.perf_counter()
        tuning = autotune(ctx, sample, params, colored=False)
        tuning["sample"] = sample_name
        log(f"[bench] autotune: {tuning['tried']} schedules on {cut} bytes in {time.perf_counter() - t:.1f} s -> "
            f"{tuning['schedule']}")
    gate = None
    if check:
        gate = parity_gate(ctx, wl, params, sample, threads)
        gate["sample"] = sample_name
        log(f"[bench] parity gate ok (schedule that is timed): {gate}")

    def step():
        ctx.bucketing(None, params, device_ptr=dptr, n_bytes=nbytes)
        ctx.kmers_merge()

    for i in range(warmup):
        step()
        log(f"[bench] warmup {i + 1}/{warmup} done")
    per_kernel = []
    t0 = time.perf_counter()
    for i in range(steps):
        step()
        per_kernel.append(ctx.timings().copy())
        log(f"[bench] step {i + 1}/{steps} done")
    dt = time.perf_counter() - t0
    ctx.free(dptr)
    cnt = ctx.counters()
    host = host_input_steps(ctx, wl, params, host_steps, cnt) if host_steps else None
    n, kw = int(cnt[5]), (1 if (wl.k <= 32 and wl.hash_kind == 0) else 2)
    ms = np.mean(np.array(per_kernel), axis=0)
    kb = kernel_bytes(cnt, kw)
    kern_ms = {"bucket_stage": ms[0], "unit_compact": ms[2], "merge": ms[3] + ms[4]}
    dom = max(kern_ms, key=lambda x: kern_ms[x])
    achieved = kb[dom] / (kern_ms[dom] * 1e-3) / 1e9
    traffic, traffic_src = measured_traffic(wl.name, scale, dom)
    ms_step = dt / steps * 1e3
    value = wl.bases / (dt / steps) / 1e9
    return {
        "metric": "input Gbases/s bucketing+kmer-merge (bit-exact); query p50/p99 ms",
        "value": round(value, 4),
        "unit": "Gbases/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8/u64 integer" if kw == 1 else "u8/u128 integer",
        "data": "synthetic (seeded generator, same shape as the named configuration)",
        "config": {"workload": wl.name, "k": wl.k, "m": wl.m, "min_multiplicity": wl.min_mult,
                   "buckets_log": wl.buckets_log, "second_buckets_log": wl.second_buckets_log,
                   "hash": "seq-hash" if wl.hash_kind == 0 else "rabin-karp128", "input_bytes": int(wl.data.size),
                   "bases": int(wl.bases)},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "algorithmic_gb": round(kb[dom] / 1e9, 3), "traffic_source": traffic_src},
        "kernels_ms": {k: round(float(v), 3) for k, v in kern_ms.items()},
        "kernels_roofline": kernel_rooflines(kb, kern_ms, wl.name, scale),
        "counters": {"records": int(cnt[1]), "stored_records": int(cnt[10]), "record_bases": int(cnt[2]),
                     "kmer_occurrences": int(cnt[3]), "split_items": int(cnt[4]), "solid_entries": int(n),
                     "merge_passes": int(cnt[7])},
        "h2d_ms": round(h2d_s * 1e3, 1),
        "tuning": tuning,
        "parity_gate": gate,
        "pcie_inclusive_gbases_s": round(wl.bases / (dt / steps + h2d_s) / 1e9, 4),
        "host_input": host,
    }


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--no-serve", action="store_true", help="skip the colored query serving measurement")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default schedule")
    ap.add_argument("--no-e2e", action="store_true", help="skip the file-to-solid-table measurement")
    ap.add_argument("--no-k63", action="store_true", help="skip the k=63 rabin-karp128 measurement (config 4)")
    ap.add_argument("--host-steps", type=int, default=2,
                    help="steps timed from the host array (chunked upload pipeline); 0 skips them")
    ap.add_argument("--serve-rate", type=float, default=20.0, help="Poisson request rate of the serving run (1/s)")
    ap.add_argument("--serve-requests", type=int, default=60)
    ap.add_argument("--schedule", type=str, default=None,
                    help="JSON object of schedule fields to use instead of the tuner's choice (config 2 line)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="run one 1 GiB device-to-device copy first (HBM counter calibration under rocprofv3 --pmc)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on exactly one GPU; sharding the path over several GPUs "
                                   "is described but not implemented"}))
        raise SystemExit(2)

    from ggcat_amd.native import GpuContext, Params

    t = time.perf_counter()
    wl = make_workload(a.config, a.scale)
    log(f"[bench] workload {wl.name}: {wl.data.size} bytes, {wl.bases} bases, k={wl.k} m={wl.m} "
        f"b={wl.buckets_log} sb={wl.second_buckets_log} ({time.perf_counter() - t:.1f} s to generate)")
    ctx = GpuContext(0)
    if a.calib_copy:
        src, dst = ctx.alloc(1 << 30), ctx.alloc(1 << 30)
        ctx.copy_device(dst, src, 1 << 30)
        ctx.free(src)
        ctx.free(dst)
        log("[bench] calibration copy (copy16_kernel): 1073741824 bytes read, 1073741824 bytes written")
    threads = min(os.cpu_count() or 1, 64)
    # the schedule is tuned and gated on the same configuration with its genome
    # scaled to a 256 MiB read set (same coverage)
    sample, sample_name = shape_sample(a.config, a.scale, wl.data.size, 256 << 20)
    out = measure(ctx, wl, a.steps, a.warmup, not a.no_autotune, not a.no_check, threads, a.scale,
                  sample=sample, sample_name=sample_name, schedule=json.loads(a.schedule) if a.schedule else None,
                  host_steps=a.host_steps)
    del sample
    n = out["counters"]["solid_entries"]
    if wl.hash_kind == 0:  # unitigs walk k-mers, which seq-hash keys are
        # unitig chains inside the sub-buckets of the last step's graph (device
        # kernels only; the sequences stay in HBM)
        t = time.perf_counter()
        n_unitigs, unitig_bytes = ctx.unitig_counts()
        k_ms = float(ctx.timings()[7])
        out["unitigs"] = {"unitigs": n_unitigs, "bytes": unitig_bytes, "kernels_ms": round(k_ms, 2),
                          "call_ms": round((time.perf_counter() - t) * 1e3, 1),
                          "solid_kmers_per_s": round(n / (k_ms / 1e3), 1) if k_ms > 0 else None}
        # maximal unitigs of the whole solid set (joined across sub-buckets)
        t = time.perf_counter()
        n_max, max_bytes = ctx.unitig_counts(joined=True)
        k_ms = float(ctx.timings()[7])
        out["maximal_unitigs"] = {"unitigs": n_max, "bytes": max_bytes, "kernels_ms": round(k_ms, 2),
                                  "call_ms": round((time.perf_counter() - t) * 1e3, 1),
                                  "solid_kmers_per_s": round(n / (k_ms / 1e3), 1) if k_ms > 0 else None}
    schedule = (out.get("tuning") or {}).get("schedule")
    if a.config == 2 and not a.no_k63:
        # the 128-bit key path at full size: k = 63, rabin-karp128, 15 Gbases
        t = time.perf_counter()
        wl4 = make_workload(4, a.scale)
        log(f"[bench] workload {wl4.name}: {wl4.data.size} bytes, {wl4.bases} bases "
            f"({time.perf_counter() - t:.1f} s to generate)")
        sample, sample_name = shape_sample(4, a.scale, wl4.data.size, 256 << 20)
        out["config4"] = measure(ctx, wl4, max(1, min(a.steps, 2)), 1, not a.no_autotune, not a.no_check, threads,
                                 a.scale, sample=sample, sample_name=sample_name, host_steps=min(a.host_steps, 1))
        del sample
        del wl4
    # the file-to-table run uses the same read set, so the same schedule
    ctx.set_tuning(**schedule) if schedule else ctx.set_tuning()
    if not a.no_e2e:
        log("[bench] end-to-end from a FASTA file ...")
        try:
            out["end_to_end"] = end_to_end(ctx, wl, schedule or {})
        except (OSError, RuntimeError, MemoryError) as e:  # e.g. no room for the temporary file
            out["end_to_end"] = {"error": f"{type(e).__name__}: {e}"}
    if not a.no_serve:
        out["query"] = serve_measurement(ctx, a.serve_rate, a.serve_requests)
    ctx.close()
    if not a.no_cpu:
        threads = min(os.cpu_count() or 1, 64)
        log(f"[bench] CPU baseline with {threads} threads ...")
        out["cpu_baseline"] = cpu_baseline(wl, threads)
        if wl.hash_kind == 0:
            log("[bench] CPU unitig baselines ...")
            out["unitigs"]["cpu_baseline"], out["maximal_unitigs"]["cpu_baseline"] = cpu_unitig_baseline(wl, threads)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
