"""Benchmark: secure-mode Ed25519 vanity-key search throughput on one MI355X.

Workload (the headline configuration): --bench style search of 2^32 ChaCha20
("secure") candidates of one master key (bytes of random.Random(2)), a seeded 6-char
case-sensitive suffix, every candidate derived, compressed and matched on the GPU.
A step is one search launch over 2^28 consecutive candidate indices; the default 16
timed steps cover the full 2^32 range once.

Reports one JSON line: keys/s (value), the search kernel's fraction of the integer
VALU multiply roofline (timed with HIP events on the library's stream), and the CPU
oracle's rate on the host cores over a bounded sample (which also checks the GPU's
public keys of that sample bit for bit).
"""
import argparse
import json
import os
import random
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "secure-mode keys/s on 1 MI355X (single vs 64-search batched); p99 time-to-hit"
ALPHABET = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/"
STEP_KEYS = 1 << 28
# MI355X integer VALU peak: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz, one 32x32->64
# multiply-add (v_mad_u64_u32) per lane per clock.
PEAK_MUL = 256 * 4 * 16 * 2.4e9


def field_mults_per_key(w: int, k: int) -> float:
    """Algorithmic field multiplications per key: 7 per extended+affine addition over
    rows - 2 rows, 3 for the last (fraction) row, 3 for the batched inversion plus 1 for
    the prefix product, and one 265-multiplication inversion shared by k keys."""
    rows = -(-251 // w)
    return 7 * (rows - 2) + 3 + 4 + 265.0 / k


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--width", type=int, default=24)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--blocks", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="only run a plain device-to-device copy of this many bytes (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_bytes:
        copy_calibration(args.copy_bytes)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (each search shards by index "
              "range across GPUs with no collectives; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import meikey_amd as M

    rng = random.Random(2)
    key = rng.randbytes(32)
    needle = "".join(rng.choice(ALPHABET) for _ in range(6)).encode()
    gpu = M.Gpu(0, width=args.width, batch=args.batch, blocks=args.blocks)
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
    mults = field_mults_per_key(args.width, args.batch)
    achieved = STEP_KEYS * mults * 64 / (avg_ms * 1e-3)
    traffic = measured_traffic(args.width, args.batch, args.blocks)
    out = {
        "metric": METRIC, "value": rate, "unit": "keys/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": elapsed * 1e3 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u32", "data": "synthetic",
        "config": {"workload": "secure-mode --bench search, 2^32 ChaCha20 candidates of one master key "
                               "(random.Random(2)), 6-char case-sensitive suffix",
                   "candidates": keys, "needle": needle.decode(), "width": args.width,
                   "keys_per_thread": args.batch, "waves_per_simd": args.blocks, "grid_waves": info.grid_waves,
                   "hits": len(all_hits)},
        "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": PEAK_MUL / 1e12, "unit": "TOP/s",
                     "frac": achieved / PEAK_MUL, "traffic": traffic,
                     "kernel": "mk_search_kernel", "kernel_ms": avg_ms,
                     "algorithmic": f"{mults:.2f} field mults/key x 64 32x32-bit products"},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(gpu, key, needle)
    gpu.close()
    print(json.dumps(out), flush=True)


def measured_traffic(width: int, batch: int, blocks: int):
    """HBM bytes per launch of the search kernel, from the committed FETCH_SIZE /
    WRITE_SIZE counter runs (profiles/mk_search_kernel_traffic.json, each scaled by the
    same counter's reading for a copy kernel of known size); None when those runs were
    made for another setting."""
    path = os.path.join(ROOT, "profiles", "mk_search_kernel_traffic.json")
    try:
        t = json.load(open(path))
    except (OSError, ValueError):
        return None
    if (t.get("width"), t.get("keys_per_thread"), t.get("waves_per_simd")) != (width, batch, blocks):
        return None
    return t["bytes_per_key"] * STEP_KEYS


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


def cpu_baseline(gpu, key, needle):
    """The CPU oracle (a scalar C port of the reference's per-key algorithm) on this
    host's cores over a bounded sample of the same workload; the GPU's public keys of
    the same sample are checked against it bit for bit."""
    from oracle import oracle as O

    cores = min(os.cpu_count() or 1, 16)
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
    return {"value": chunks * per / dt, "unit": "keys/s", "cores": cores, "kind": "port",
            "sample": f"{chunks} x {per} consecutive candidates spread over the 2^32 range, derive + suffix match "
                      f"with the C oracle; GPU public keys of the same 2^20 candidates compared: "
                      f"{'all equal' if mismatches == 0 else f'{mismatches} chunks differ'}",
            "checked_keys": chunks * per, "mismatched_chunks": mismatches}


if __name__ == "__main__":
    main()
