"""Host side of the query path: output formatting of per-query counts exactly
as the reference writes them, and a continuous-batching serving loop over the
C ABI.

Formats:
* uncolored CSV (crates/querier/src/pipeline/counters_sorting.rs:235-270):
  header ``query_index,matched_kmers,query_kmers,match_percentage``, one row
  per query in input order, 0-based index, ratio printed with two decimals;
* colored JSON lines (crates/querier/src/pipeline/colored_query_output.rs:
  255-300): one line per query with at least one matched k-mer,
  ``{"query_index":Q, "matches":{"c": v,...}}`` with colors ascending and
  v = count / query_kmers printed with two decimals; queries without any
  color count are omitted.  Color names replace the numbers with the
  names format (``JsonLinesWithNames``).
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np


def _ratio(num: int, den: int) -> str:
    # Rust's "{:.2}" of an f64: the exact binary value rounded to 2 decimals,
    # as Python's format does; 0/0 prints NaN like the reference's f64 division
    if den == 0:
        return "NaN" if num == 0 else "inf"
    return f"{num / den:.2f}"


def format_csv(query_kmers: np.ndarray, matched: np.ndarray) -> str:
    lines = ["query_index,matched_kmers,query_kmers,match_percentage"]
    for i, (qk, mt) in enumerate(zip(query_kmers.tolist(), matched.tolist())):
        lines.append(f"{i},{mt},{qk},{_ratio(mt, qk)}")
    return "\n".join(lines) + "\n"


def format_jsonl(query_kmers: np.ndarray, color_counts: np.ndarray, color_names: list[str] | None = None) -> str:
    out = []
    for q in range(color_counts.shape[0]):
        cols = np.flatnonzero(color_counts[q])
        if cols.size == 0:
            continue
        qk = int(query_kmers[q])
        items = []
        for c in cols.tolist():
            name = color_names[c] if color_names is not None else str(c)
            items.append(f'"{name}": {_ratio(int(color_counts[q, c]), qk)}')
        out.append('{"query_index":%d, "matches":{%s}}' % (q, ",".join(items)))
    return "".join(line + "\n" for line in out)


@dataclass
class ServeReport:
    latencies_ms: list[float] = field(default_factory=list)
    queries: int = 0
    kmers: int = 0
    wall_s: float = 0.0
    max_outstanding: int = 0
    max_admitted: int = 0

    def percentile(self, p: float) -> float:
        return float(np.percentile(np.array(self.latencies_ms), p)) if self.latencies_ms else 0.0


def serve_poisson(ctx, requests: list[np.ndarray], rate_hz: float, seed: int = 5, results: dict | None = None,
                  poll_sleep_s: float = 0.0002) -> ServeReport:
    """Submits `requests` (joined query batches) with exponential inter-arrival
    times of mean 1/rate_hz, polling for finished requests in between; the
    latency of a request is submission to the poll that returns it.  With
    `results`, stores (query_kmers, matched, color_counts) per request index."""
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate_hz, len(requests)))
    rep = ServeReport()
    sent_at: dict[int, float] = {}
    which: dict[int, int] = {}
    nq = [int(np.count_nonzero(r == 10)) for r in requests]
    t0 = time.perf_counter()
    nxt = 0
    done = 0
    while done < len(requests):
        now = time.perf_counter() - t0
        while nxt < len(requests) and arrivals[nxt] <= now:
            t = ctx.serve_submit(requests[nxt])
            sent_at[t] = time.perf_counter()
            which[t] = nxt
            nxt += 1
        out, adm = ctx.serve_stats()
        rep.max_outstanding = max(rep.max_outstanding, out)
        rep.max_admitted = max(rep.max_admitted, adm)
        t = ctx.serve_poll(False)
        if t is None:
            if nxt < len(requests):
                time.sleep(min(poll_sleep_s, max(0.0, arrivals[nxt] - (time.perf_counter() - t0))))
            else:
                time.sleep(poll_sleep_s)
            continue
        end = time.perf_counter()
        i = which[t]
        res = ctx.serve_fetch(t, nq[i])
        rep.latencies_ms.append((end - sent_at[t]) * 1e3)
        rep.queries += nq[i]
        rep.kmers += int(res[0].sum())
        if results is not None:
            results[i] = res
        done += 1
    rep.wall_s = time.perf_counter() - t0
    return rep
