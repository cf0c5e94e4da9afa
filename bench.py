"""Benchmark: secure-mode Ed25519 vanity-key search throughput on one MI355X.

Workload (the headline configuration): --bench style search of 2^32 ChaCha20
("secure") candidates of one master key (bytes of random.Random(2)), a seeded 6-char
case-sensitive suffix, every candidate derived, compressed and matched on the GPU.
A step is one search launch over 2^29 consecutive candidate indices (about 270 ms, the
order of the reference's 200 ms launch target); the default 8 timed steps cover the full
2^32 range once.

Reports one JSON line: keys/s (value), the search kernel's fraction of the integer
VALU multiply roofline (timed with HIP events on the library's stream), the 64-search
batched launch against the single search and against one search per launch, the
serving run (Poisson arrivals, p50/p99 time to first hit), and the CPU baseline: the
restatement of the reference's SIMD CPU search (AVX-512 IFMA or AVX2, oracle/simd) timed
for 10 s on pinned host threads, with the reference kernel's host build and the scalar C
oracle beside it (the oracle also checks the GPU's public keys of its sample, and every
key the server returned, bit for bit).
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "secure-mode keys/s on 1 MI355X (single vs 64-search batched); p99 time-to-hit"
STEP_KEYS = 1 << 29
ctypes_u32, ctypes_u64, ctypes_buf = ctypes.c_uint32, ctypes.c_uint64, ctypes.create_string_buffer
# MI355X integer VALU peak: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz, one 32x32->64
# multiply-add (v_mad_u64_u32) per lane per clock. Measured: v_mad_u64_u32 issues one
# wave64 instruction per SIMD per 4 clocks and never pairs with another VALU instruction;
# only plain 32-bit operations dual-issue (meikey_amd/mk_valu_probe).
PEAK_MUL = 256 * 4 * 16 * 2.4e9


def field_mults_per_key(w: int, k: int, pool: int = 0) -> float:
    """Algorithmic field multiplications per key: 7 per extended+affine addition over
    rows - 2 rows, 3 for the last (fraction) row, 3 for the batched inversion plus 1 for
    the prefix product, and one 265-multiplication inversion shared by k keys (with a
    pooled inversion, by the pool x k keys of a lane's chain, which adds 3 (pool - 1)
    multiplications)."""
    rows = -(-251 // w)
    inv = 265.0 / k if not pool else (265.0 + 3 * (pool - 1)) / (pool * k)
    return 7 * (rows - 2) + 3 + 4 + inv


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--width", type=int, default=24)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--blocks", type=int, default=2)
    ap.add_argument("--limbs", type=int, default=None,
                    help="field / table layout: 10 (packed ten-limb), 11 (ten-limb entries), 8 (eight words), "
                         "26 / 27 (Karatsuba ten-limb, packed / limb entries); default: the library's")
    ap.add_argument("--pool", type=int, default=None,
                    help="4 or 8: workgroups of that many waves share one inversion (default: the library's)")
    ap.add_argument("--lockstep", type=int, default=None,
                    help="1: one workgroup per CU adding each comb row together (default: the library's)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="timed length of the SIMD CPU baseline")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-search batched leg")
    ap.add_argument("--serve-seconds", type=float, default=60.0, help="serving run length (0 = skip)")
    ap.add_argument("--no-low", action="store_true", help="skip the low-security (counter seed) leg")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="only run a plain device-to-device copy of this many bytes (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_bytes:
        copy_calibration(args.copy_bytes)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (this build runs on one GPU; a "
              "search would shard by disjoint index ranges across GPUs, with no collectives)", file=sys.stderr)
        sys.exit(2)

    import meikey_amd as M
    from meikey_amd import workloads as W

    key, needle = W.single()
    needle = needle.encode()
    if args.lockstep is None:
        args.lockstep = M.default_options().lockstep
    if args.pool is None:
        args.pool = M.default_options().pool
    if args.limbs is None:
        args.limbs = M.default_options().limbs
    gpu = M.Gpu(0, width=args.width, batch=args.batch, blocks=args.blocks, limbs=args.limbs, lockstep=args.lockstep,
                pool=args.pool)
    info = gpu.info()
    print(f"# device {info.name.decode()} CUs={info.compute_units} grid={info.grid_waves} waves "
          f"W={args.width} batch={args.batch} table={info.table_bytes / 2**20:.1f} MiB needle={needle.decode()}",
          file=sys.stderr, flush=True)

    def step(first):
        hits, tries = gpu.batch_run([{"key": key, "seed_source": M.SEED_CHACHA20, "needle": needle, "fold": False,
                                      "mode": M.MODE_SUFFIX, "first": first, "count": STEP_KEYS}])
        return hits, tries[0], gpu.last_timing()[0]

    for i in range(args.warmup):
        step((1 << 40) + i * STEP_KEYS)
    all_hits, keys, kernel_ms = [], 0, []
    t0 = time.perf_counter()
    for i in range(args.steps):
        hits, tries, ms = step(i * STEP_KEYS)
        all_hits += hits
        keys += tries
        kernel_ms.append(ms)
        print(f"# step {i} kernel {ms:.1f} ms", file=sys.stderr, flush=True)
    elapsed = time.perf_counter() - t0
    rate = keys / elapsed
    avg_ms = sum(kernel_ms) / len(kernel_ms)
    # Algorithmic work per key is the canonical W=16, K=256 figure (106.04 field
    # multiplications x 64 products = 6,787 32x32-bit products), whatever width and batch
    # the run uses, so the achieved rate is proportional to keys/s; the run's own count
    # (fewer multiplications at larger widths) is reported beside it.
    mults = field_mults_per_key(16, 256)
    run_mults = field_mults_per_key(args.width, args.batch, args.pool)
    achieved = STEP_KEYS * mults * 64 / (avg_ms * 1e-3)
    achieved_run = STEP_KEYS * run_mults * 64 / (avg_ms * 1e-3)
    traffic = measured_traffic(args.width, args.batch, args.blocks)
    issue = (profiled(args.width, args.batch, args.blocks) or {}).get("valu_issue") or {}
    # the same peak at the clock the box sustains through a full-chip launch (the counter
    # run's GRBM_GUI_ACTIVE over its time)
    peak_clock = 256 * 4 * 16 * issue["clock_ghz"] * 1e9 if issue.get("clock_ghz") else None
    out = {
        "metric": METRIC, "value": rate, "unit": "keys/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": elapsed * 1e3 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u32", "data": "synthetic",
        "config": {"workload": f"secure-mode --bench search, {args.steps} launches of 2^29 consecutive ChaCha20 "
                               f"candidates ({keys} in all; 8 launches cover the 2^32 range once) of one master key "
                               f"(random.Random(2)), 6-char case-sensitive suffix",
                   "candidates": keys,
                   "inversion": (f"one per {args.batch} keys of a lane, pooled over the {args.pool} waves of a "
                                 f"workgroup" if args.pool else f"one per {args.batch} keys of a lane"), "needle": needle.decode(), "width": args.width,
                   "keys_per_thread": args.batch, "waves_per_simd": args.blocks, "grid_waves": info.grid_waves,
                   "lockstep": args.lockstep, "pool": args.pool, "limbs": args.limbs,
                   "hits": len(all_hits)},
        "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": PEAK_MUL / 1e12, "unit": "TOP/s",
                     "frac": achieved / PEAK_MUL, "traffic": traffic,
                     "kernel": "mk_search_kernel", "kernel_ms": avg_ms,
                     "algorithmic": f"{mults:.2f} field mults/key (W=16, K=256) x 64 32x32-bit products",
                     "achieved_run_mults": achieved_run / 1e12,
                     "run_mults_per_key": run_mults,
                     # the multiplications this run does (W and K as run), against the
                     # same 39.3 T peak and against the peak at the sustained clock
                     "frac_run_mults": achieved_run / PEAK_MUL,
                     "peak_at_sustained_clock": peak_clock / 1e12 if peak_clock else None,
                     "sustained_clock_ghz": issue.get("clock_ghz"),
                     "frac_at_sustained_clock": achieved / peak_clock if peak_clock else None,
                     "frac_run_mults_at_sustained_clock": achieved_run / peak_clock if peak_clock else None,
                     # integer VALU is the bound (no dense contraction for MFMA); the HBM
                     # side of the same launches, from the committed traffic counters
                     "hbm": ({"achieved": traffic / (avg_ms * 1e-3) / 1e12, "peak": 8.0, "unit": "TB/s",
                              "frac": traffic / (avg_ms * 1e-3) / 8e12} if traffic else None),
                     # the bound itself: how often the SIMDs issue a VALU instruction, from
                     # the committed SQ counter run of the same launch
                     "valu_issue": issue or None},
    }
    # The headline pattern's time to a first hit. Each candidate matches independently with
    # the pattern's probability p (Pattern::probability, src/model/pattern.rs), so the tries
    # to the first hit are geometric: mean 1/p, q-quantile ln(1/(1-q))/p. At the measured
    # single-search rate these are times; they are derived, not a sample of searches.
    pat = M.Pattern(needle.decode(), True, M.MODE_SUFFIX)
    mean_tries = pat.expected_tries
    out["time_to_hit"] = {
        "pattern": needle.decode(), "mode": "suffix, case-sensitive", "expected_tries": mean_tries,
        "mean_s": mean_tries / rate, "p50_s": math.log(2) * mean_tries / rate,
        "p99_s": math.log(100) * mean_tries / rate,
        "basis": "geometric tries to the first hit (probability 1/expected_tries per candidate) at the measured "
                 "single-search keys/s; the serving leg measures times to first hit of a request mix"}
    print(f"# headline pattern: mean {mean_tries / rate:.1f} s to a first hit, p99 "
          f"{math.log(100) * mean_tries / rate:.1f} s at {rate / 1e9:.3f} G keys/s", file=sys.stderr, flush=True)
    if not args.no_batch:
        out["batched"] = batched_leg(gpu, rate)
        out["batched"]["kernel_vs_single_kernel"] = out["batched"]["kernel_keys_per_s"] / (STEP_KEYS / (avg_ms * 1e-3))
    low = None
    if not args.no_low:
        out["low_security"], low = low_security_leg(gpu, args)
        out["low_security"]["vs_secure"] = out["low_security"]["value"] / rate
    served = []
    if args.serve_seconds > 0:
        out["serving"], served = serving_leg(gpu, args.serve_seconds)
        # the same arrivals served in steps half as long (64-key claim units): lower latency
        out["serving_low_latency"], _ = serving_leg(gpu, min(args.serve_seconds, 20.0), 1 << 23, 64)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(gpu, key, needle, served, low, args.cpu_seconds)
    gpu.close()
    print(json.dumps(out), flush=True)


def batched_leg(gpu, single_rate):
    """64 searches of mixed kinds (substring either case, suffix, regex DFA), 2^26
    candidates each, in one launch; then the same 64 searches one launch each. The
    untimed warm-up is the same batch, so the timed call finds the host hit buffers
    (about 600k hits) already in memory, as every later call of a long-running caller does:
    the records land in a numpy array the caller keeps (batch_run's out)."""
    import numpy as np
    import meikey_amd as M
    from meikey_amd import workloads as W
    specs, pats, keys = W.batched(1 << 26)
    buf = np.zeros(1 << 20, dtype=M.hit_dtype())
    gpu.batch_run(specs, out=buf)
    t0 = time.perf_counter()
    hits, tries = gpu.batch_run(specs, out=buf)
    dt = time.perf_counter() - t0
    batched_ms = gpu.last_timing()[0]
    total = sum(tries)
    t0 = time.perf_counter()
    single_hits = []
    single_ms = 0.0
    for s in specs:
        h, _ = gpu.batch_run([s], as_array=True)
        single_ms += gpu.last_timing()[0]
        single_hits.append(h["index"])
    dt1 = time.perf_counter() - t0
    by = [hits["index"][hits["search"] == i] for i in range(len(specs))]
    mismatched = sum(1 for a, b in zip(by, single_hits) if a.tolist() != b.tolist())
    rate = total / dt
    print(f"# batched 64 searches: {rate / 1e9:.3f} G keys/s, one per launch {total / dt1 / 1e9:.3f} G keys/s, "
          f"{len(hits)} hits, {mismatched} searches differ", file=sys.stderr, flush=True)
    return {"value": rate, "unit": "keys/s", "searches": len(specs), "candidates": total,
            "vs_single": rate / single_rate, "one_per_launch": total / dt1,
            "vs_one_per_launch": dt1 / dt, "hits": len(hits), "mismatched_searches": mismatched,
            "kernel_keys_per_s": total / (batched_ms * 1e-3),
            "one_per_launch_kernel_keys_per_s": total / (single_ms * 1e-3),
            "kinds": "16 substring case-insensitive 3-7, 16 substring case-sensitive 3-7, 16 suffix 3-6, "
                     "16 regex (<= 64 DFA states)"}


def low_security_leg(gpu, args):
    """Low-security mode: counter seeds of the master key of random.Random(3), an 8-char
    case-insensitive substring, 2^32 candidates in 8 launches. Beside it the same search
    with a cache-resident comb (W = 6: 42 rows x 33 entries x 128 B = 177 KB, gathered
    through L2) and with the comb table in LDS (W = 5 and W = 6), to show the table-size /
    field-work trade."""
    import random
    import meikey_amd as M
    from meikey_amd import workloads as W
    rng = random.Random(3)
    key = rng.randbytes(32)
    needle = "".join(rng.choice(W.ALPHABET) for _ in range(8)).lower().encode()

    def run(g, launches, step):
        spec = {"key": key, "seed_source": M.SEED_COUNTER, "needle": needle, "fold": True, "mode": M.MODE_CONTAINS}
        g.batch_run([dict(spec, first=1 << 40, count=step)])
        hits, ms, t0 = [], 0.0, time.perf_counter()
        for i in range(launches):
            h, _ = g.batch_run([dict(spec, first=i * step, count=step)])
            hits += h
            ms += g.last_timing()[0]
        return launches * step / (time.perf_counter() - t0), launches * step / (ms * 1e-3), hits

    launches = (1 << 32) // STEP_KEYS
    rate, kernel_rate, hits = run(gpu, launches, STEP_KEYS)
    small = M.Gpu(0, width=6, batch=args.batch, blocks=args.blocks)
    small_bytes = small.info().table_bytes
    small_rate, small_kernel, _ = run(small, 2, 1 << 26)
    small.close()
    # the comb table in LDS (each CU's copy, read without HBM / L2 gathers): W = 5 with
    # 128-byte entries and W = 6 with 96-byte entries, 32 keys per lane
    lds = {}
    for w, limbs in ((5, M.LIMBS_R25U), (6, M.LIMBS_R25)):
        g = M.Gpu(0, width=w, batch=32, blocks=2, limbs=limbs, lds_table=1)
        lds_bytes = g.info().table_bytes
        r, kr, _ = run(g, 2, 1 << 26)
        g.close()
        lds[f"w{w}"] = {"keys_per_s": r, "kernel_keys_per_s": kr, "table_bytes": lds_bytes, "limbs": limbs,
                        "candidates": 2 << 26}
    print(f"# low-security: {rate / 1e9:.3f} G keys/s (W={args.width}), {small_rate / 1e9:.3f} G keys/s (W=6 "
          f"cache-resident table), {lds['w5']['keys_per_s'] / 1e9:.3f} / {lds['w6']['keys_per_s'] / 1e9:.3f} G keys/s "
          f"(W=5 / W=6 table in LDS), {len(hits)} hits", file=sys.stderr, flush=True)
    return ({"value": rate, "unit": "keys/s", "candidates": launches * STEP_KEYS, "needle": needle.decode(),
             "kernel_keys_per_s": kernel_rate, "hits": len(hits), "vs_secure": None,
             "w6_table": {"keys_per_s": small_rate, "kernel_keys_per_s": small_kernel, "table_bytes": small_bytes,
                          "candidates": 2 << 26},
             "lds_table": lds},
            (key, needle))


def percentile(xs, q):
    xs = sorted(xs)
    if not xs:
        return None
    k = min(len(xs) - 1, max(0, int(round(q * (len(xs) - 1)))))
    return xs[k]


def serving_leg(gpu, seconds, step_keys=1 << 24, unit_keys=128):
    """Poisson arrivals at 50/s for `seconds`, pattern length 3-6, kinds mixed 25% each,
    every request stopping at its first hit. Time to first hit = done - submit. Requests
    unfinished when the arrivals stop are cancelled and counted with their elapsed time,
    a lower bound (censored); p50/p99 are over all submitted requests."""
    import meikey_amd as M
    from meikey_amd import workloads as W
    reqs = W.serving(50.0, seconds)
    srv = M.Server(gpu, max_slots=4096, use_graph=True, dfa_budget=256 << 20, step_keys=step_keys,
                   active_per_step=16, unit_keys=unit_keys)
    ids, results, i = {}, {}, 0
    t0 = srv.now()
    steps = 0
    while True:
        now = srv.now() - t0
        while i < len(reqs) and reqs[i].arrival_s <= now:
            ids[srv.submit(reqs[i].spec())] = i
            i += 1
        if now >= seconds:
            break
        srv.step()
        steps += 1
        for r in srv.poll():
            results[ids[r.request]] = r
    end = srv.now() - t0
    for rid, k in ids.items():
        if k not in results:
            srv.cancel(rid)
    for r in srv.poll():
        results[ids[r.request]] = r
    _, step_keys = srv.stats()
    srv.close()
    found = [r for r in results.values() if r.status == 1]
    ttfh = [r.done_s - r.submit_s for r in found]
    censored = [end - (r.submit_s - t0) for r in results.values() if r.status != 1]
    everything = ttfh + censored
    tries = sum(r.tries for r in results.values())
    p99 = percentile(everything, 0.99)
    out = {"requests": len(ids), "found": len(found), "unfinished": len(censored), "seconds": end,
           "steps": steps, "step_keys": step_keys, "keys_per_s": tries / end,
           "p50_time_to_first_hit_s": percentile(everything, 0.5), "p99_time_to_first_hit_s": p99,
           "p99_censored": bool(censored) and p99 is not None and p99 >= min(censored),
           "p50_found_only_s": percentile(ttfh, 0.5), "p99_found_only_s": percentile(ttfh, 0.99),
           "rate_per_s": 50.0,
           "policy": f"least attained service, 16 requests per step, {step_keys} candidates per step claimed in "
                     f"units of {unit_keys} keys per lane, two steps in flight on two streams serving disjoint "
                     "requests",
           "censoring": "requests unfinished at the end count with their elapsed time (lower bound)"}
    print(f"# serving ({unit_keys}-key units): {len(found)}/{len(ids)} found, p50 {out['p50_time_to_first_hit_s']:.3f} s, "
          f"p99 {p99:.3f} s, {tries / end / 1e9:.3f} G keys/s", file=sys.stderr, flush=True)
    served = [(reqs[k], r) for k, r in results.items() if r.status == 1]
    return out, served


def profiled(width: int, batch: int, blocks: int):
    """The committed counter runs of the search kernel (profiles/mk_search_kernel_traffic.json):
    FETCH_SIZE / WRITE_SIZE, each scaled by the same counter's reading for a copy kernel
    of known size, and the VALU issue rate of the SQ counters. None when those runs were
    made for another setting."""
    path = os.path.join(ROOT, "profiles", "mk_search_kernel_traffic.json")
    try:
        t = json.load(open(path))
    except (OSError, ValueError):
        return None
    if (t.get("width"), t.get("keys_per_thread"), t.get("waves_per_simd")) != (width, batch, blocks):
        return None
    return t


def measured_traffic(width: int, batch: int, blocks: int):
    """HBM bytes per launch of the search kernel, from the committed counter runs."""
    t = profiled(width, batch, blocks)
    return t["bytes_per_key"] * STEP_KEYS if t else None


def copy_calibration(nbytes: int) -> None:
    """A copy kernel of known size: read nbytes and write nbytes, three times. Its
    FETCH_SIZE / WRITE_SIZE counters give the scale factor for the search kernel's."""
    import torch
    x = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    x.fill_(1)
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": nbytes, "dispatches": 3}), flush=True)


def reference_cpu_leg(key, needle, cores):
    """The reference's own search kernel compiled as host C++ (oracle/_ref, from
    kernel.cu), run on `cores` host threads over a bounded sample of the headline
    workload, with the reference's default W = 16 comb table (built on the device and
    repacked into the reference layout). Its public keys of sampled indices are compared
    with the GPU's. None where the reference build is absent."""
    import struct
    import threading
    import meikey_amd as M
    from oracle import oracle as O
    R = O.ref()
    if R is None:
        return None
    g16 = M.Gpu(0, width=16, batch=8, blocks=1)
    try:
        info = g16.info()
        import numpy as np
        rows = [g16.table_words(r, 0, info.cols, as_array=True) for r in range(info.rows)]
        table = O.ref_table_from_device(np.concatenate(rows))
        kw = (ctypes_u32 * 8)(*struct.unpack("<8I", key))
        batch = R.meikey_test_batch()
        wrong = 0
        for first in (0, 12345, (1 << 31) + 7, (1 << 32) - batch):
            out = ctypes_buf(32 * batch)
            R.meikey_test_derive_batch(table, 16, kw, 0, first, out)
            if out.raw != g16.derive(key, 0, first, batch):
                wrong += 1
        nthreads, iters = 128, 16  # one call covers 128 x 16 x 32 = 65536 candidates
        per_call = nthreads * iters * batch
        calls_per_core = 4
        def work(c):
            for j in range(calls_per_core):
                res = (ctypes_u64 * 2)()
                base = (c * calls_per_core + j) * (1 << 26)
                R.meikey_test_search(nthreads, table, 16, kw, 0, needle, len(needle), 0, 1, base, iters, 0, res)
        ts = [threading.Thread(target=work, args=(c,)) for c in range(cores)]
        t0 = time.perf_counter()
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        dt = time.perf_counter() - t0
        keys = cores * calls_per_core * per_call
        return {"value": keys / dt, "unit": "keys/s", "cores": cores, "kind": "reference",
                "sample": f"{cores} threads x {calls_per_core} runs of the reference kernel's host build "
                          f"(meikey_test_search, 128 threads x 16 iterations x 32 keys = {per_call} candidates "
                          f"each, W = 16 table), the headline key and suffix",
                "reference_keys_checked": 4 * batch, "reference_keys_differing_from_gpu": wrong}
    finally:
        g16.close()


# Host threads for the CPU legs: every CPU of the affinity mask, at most the 16 CPUs that
# a one-GPU share of the host provides.
CPU_SHARE = 16


def simd_cpu_leg(gpu, key, needle, seconds, low=None, side_seconds=5.0):
    """The restatement of the reference's SIMD CPU search path (oracle/libmeikey_simd.so:
    AVX-512 IFMA 5 x 51-bit lanes, or AVX2 where the CPU has no IFMA; 8-way SHA-512; W = 12
    comb; 256-key batch inversion; unsigned-first matching), timed on pinned threads over
    the headline workload (`seconds`), and for `side_seconds` each over the low-security
    workload (counter seeds, 8-char case-insensitive substring) and the 64 mixed searches
    of the batched workload, one after another. Its public keys of 2^16 sampled candidates
    are compared with the GPU's first. None where the CPU has neither instruction set."""
    from oracle import simd as S
    try:
        S.lib()
    except RuntimeError as e:
        print(f"cpu baseline: SIMD leg skipped ({e})", file=sys.stderr, flush=True)
        return None
    affinity = len(os.sched_getaffinity(0))
    threads = max(1, min(affinity, CPU_SHARE))
    model = S.cpu_model()
    isa = {512: "AVX-512 IFMA", 2: "AVX2"}[S.isa()]
    print(f"cpu baseline: {model}, {isa}, {threads} pinned threads ({affinity} CPUs in the affinity mask), "
          f"{seconds:.0f} s", file=sys.stderr, flush=True)
    checked = wrong = 0
    for first in (0, 12345, (1 << 31) + 7, (1 << 32) - 4096):
        n = 1 << 14
        a, b = S.derive(key, 0, first, n), gpu.derive(key, 0, first, n)
        checked += n
        wrong += sum(1 for i in range(n) if a[32 * i:32 * i + 32] != b[32 * i:32 * i + 32])
    r = S.bench(key, needle, True, 1, seconds, threads=threads)
    out = {"value": r["keys_per_s"], "unit": "keys/s", "cores": threads, "kind": "port",
           "sample": f"{r['keys']} consecutive candidates of the headline key and suffix in {r['seconds']:.2f} s "
                     f"on {threads} pinned threads, each from its own 2^40-aligned start (restatement of the "
                     f"reference's SIMD CPU search, {isa} engine, W = 12 comb, 256-key inversion, unsigned-first "
                     f"matching)",
           "cpu_model": model, "isa": isa, "affinity_cpus": affinity,
           "per_thread": r["keys_per_s"] / threads, "hits": r["hits"],
           "simd_keys_checked_against_gpu": checked, "simd_keys_differing_from_gpu": wrong}
    if side_seconds > 0:
        import meikey_amd as M
        from meikey_amd import workloads as W
        if low is not None:
            lkey, lneedle = low
            lr = S.bench(lkey, lneedle, False, M.MODE_CONTAINS, side_seconds, threads=threads, source=M.SEED_COUNTER)
            out["low_security"] = {"keys_per_s": lr["keys_per_s"], "keys": lr["keys"], "seconds": lr["seconds"],
                                   "hits": lr["hits"], "per_thread": lr["keys_per_s"] / threads}
        specs, _, _ = W.batched(1 << 26)
        keys = hits = 0
        t = 0.0
        for sp in specs:
            br = S.bench(sp["key"], sp["needle"], not sp["fold"], sp["mode"], side_seconds / len(specs),
                         threads=threads)
            keys += br["keys"]
            hits += br["hits"]
            t += br["seconds"]
        out["batched"] = {"keys_per_s": keys / t, "keys": keys, "seconds": t, "hits": hits, "searches": len(specs),
                          "per_thread": keys / t / threads,
                          "sample": f"each of the 64 searches for {side_seconds / len(specs):.3f} s on all threads, "
                                    f"one after another"}
    return out


def cpu_baseline(gpu, key, needle, served=(), low=None, simd_seconds=10.0):
    """CPU baselines on this host's cores over bounded samples of the same workload: the
    restatement of the reference's SIMD CPU search (the line reported), the reference's
    GPU kernel compiled for the host (where oracle/_ref was built) and the scalar C
    oracle. The GPU's public keys of the oracle's sample, every key the server returned
    and (low-security leg) the keys of every 2^16-th counter-seed index are checked
    against the oracle."""
    from oracle import oracle as O

    cores = max(1, min(len(os.sched_getaffinity(0)), CPU_SHARE))
    chunks, per = 16, 1 << 16  # 2^20 keys spread over the 2^32 range
    firsts = [c * (1 << 28) + 12345 for c in range(chunks)]
    t0 = time.perf_counter()
    for f in firsts:
        O.search(key, 0, f, per, needle, True, 1, threads=cores)
    dt = time.perf_counter() - t0
    mismatches = 0
    for f in firsts:
        if gpu.derive(key, 0, f, per) != O.derive(key, 0, f, per, threads=cores):
            mismatches += 1
    bad_served = 0
    for req, r in served:
        pk = O.derive(req.key, 0, r.index, 1)
        if pk != r.pk or O.seed(req.key, 0, r.index) != r.seed or not req.pattern.is_match(O.public_b64(pk)):
            bad_served += 1
    low_checked = low_wrong = 0
    if low is not None:
        # every 2^16-th counter-seed index of the 2^32 range: 65,536 keys, the GPU's one by
        # one against the oracle's (computed on the host threads)
        from concurrent.futures import ThreadPoolExecutor
        lkey, _ = low
        idx = list(range(0, 1 << 32, 1 << 16))
        gpu_pks = [gpu.derive(lkey, 1, i, 1) for i in idx]
        with ThreadPoolExecutor(cores) as ex:
            cpu_pks = list(ex.map(lambda i: O.derive(lkey, 1, i, 1), idx, chunksize=256))
        low_checked = len(idx)
        low_wrong = sum(1 for a, b in zip(gpu_pks, cpu_pks) if a != b)
    port = {"value": chunks * per / dt, "unit": "keys/s", "cores": cores, "kind": "port",
            "sample": f"{chunks} x {per} consecutive candidates spread over the 2^32 range, derive + suffix match "
                      f"with the C oracle; GPU public keys of the same 2^20 candidates compared: "
                      f"{'all equal' if mismatches == 0 else f'{mismatches} chunks differ'}"}
    checks = {"served_keys_checked": len(served), "served_keys_wrong": bad_served,
              "low_security_keys_checked": low_checked, "low_security_keys_wrong": low_wrong,
              "checked_keys": chunks * per, "mismatched_chunks": mismatches}
    # the SIMD restatement of the reference's CPU path is the baseline; the reference
    # kernel's host build and the scalar oracle are reported beside it
    simd = simd_cpu_leg(gpu, key, needle, simd_seconds, low)
    ref = reference_cpu_leg(key, needle, cores)
    out = dict(simd) if simd else dict(ref) if ref else dict(port)
    if simd and ref:
        out["reference_kernel_host_build"] = ref
    out["scalar_oracle"] = port
    out.update(checks)
    return out


if __name__ == "__main__":
    main()
