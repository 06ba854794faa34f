"""Full-size parity of the benchmark workload (config 2: 4.0 Gbases, k=31,
m=15, s=2): the GPU's solid table against the oracle's, compared through
order-free digests (entry count, xor and sum of a 64-bit mix of sub-bucket,
key, multiplicity and flags, multiplicity sum).  Not collected by pytest: it
needs a GPU, ~1.5 minutes and ~90 GB of host memory for the oracle.  An
optional JSON schedule (ggcat_gpu_set_tuning fields, e.g. the one bench.py's
tuner picks) runs the GPU side under that schedule.

    python tests/full_size_parity.py ['{"unit_bytes": 1000, ...}']
"""
import json
import os
import sys
import time

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import numpy as np
from ggcat_amd import synth
from ggcat_amd.native import GpuContext, Params
from oracle import binding as O

def digest(sub, lo, hi, mult, flags):
    with np.errstate(over="ignore"):
        x = (lo.astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15)) ^ (hi.astype(np.uint64) * np.uint64(0xC2B2AE3D27D4EB4F))
        x ^= (sub.astype(np.uint64) << np.uint64(40)) ^ (mult.astype(np.uint64) << np.uint64(8)) ^ flags.astype(np.uint64)
        x ^= x >> np.uint64(31); x *= np.uint64(0xBF58476D1CE4E5B9); x ^= x >> np.uint64(29)
        return int(lo.size), int(np.bitwise_xor.reduce(x)), int(x.sum(dtype=np.uint64)), int(mult.astype(np.uint64).sum())

wl = synth.config2(1.0)
P = Params(31, 15, wl.buckets_log, wl.second_buckets_log, 0, 0, 2)
ctx = GpuContext(0)
schedule = json.loads(sys.argv[1]) if len(sys.argv) > 1 else {}
if schedule:
    ctx.set_tuning(**schedule)
print("[parity] schedule", schedule or "default", flush=True)
p, n = ctx.upload(wl.data)
ctx.bucketing(None, P, device_ptr=p, n_bytes=n)
ctx.kmers_merge()
c = ctx.counters()
print(f"[parity] gpu records {int(c[1])}, items in partitioned passes {int(c[4])}, merge passes {int(c[7])}",
      flush=True)
s = ctx.solid()
g = digest(s.sub, s.keys[:, 0], np.zeros_like(s.keys[:, 0]), s.mult, s.flags)
print("[parity] gpu", g, flush=True)
del s
ctx.close()
t = time.time()
o = O.bucketing(wl.data, 31, 15, wl.buckets_log, wl.second_buckets_log, threads=16)
print(f"[parity] oracle bucketing {time.time() - t:.1f} s", flush=True)
o.merge(31, wl.buckets_log, wl.second_buckets_log, 0, 2, 16)
print(f"[parity] oracle merge done {time.time() - t:.1f} s", flush=True)
os_ = o.solid()
e = digest(os_["sub"], os_["key_lo"], os_["key_hi"], os_["mult"], os_["flags"])
print("[parity] oracle", e, flush=True)
print(json.dumps({"equal": g == e, "gpu": g, "oracle": e, "schedule": schedule, "split_items": int(c[4])}),
      flush=True)
sys.exit(0 if g == e else 1)
