"""Autotuner for the sort and scan tile shapes.

Sweeps the onesweep sort over radix bits {4, 8, 9} x keys/thread {8, 15, 23}
x workgroup {256, 512, 1024} (combinations that are not built or whose tile
does not fit LDS are skipped) for u32 and u64 (key, index) pairs and for u64
keys alone (36-bit keys, the edit pipeline's packed record keys), and the
scan over its tile shapes, timing each with hipEvents on the workload it is
asked for; the best shape per row is written to voxel_amd/tuning.json, which
`voxel_amd` applies at import.

    python -m voxel_amd.autotune [--sort-keys 26] [--batched]

--batched also times the sweep on the batched multi-scene edit (the sort of
(scene, Morton) keys inside one pipeline) and lets that workload decide the
u64 row.
"""
from __future__ import annotations

import argparse
import itertools
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
TUNING_PATH = os.path.join(HERE, "tuning.json")

RADIX_BITS = (4, 8, 9)
KEYS_PER_THREAD = (8, 15, 23)
BLOCKS = (256, 512, 1024)
SCAN_SHAPES = ((256, 4), (256, 8), (256, 16), (512, 4), (512, 8), (1024, 4))


def load() -> dict:
    if not os.path.exists(TUNING_PATH):
        return {}
    with open(TUNING_PATH) as f:
        return json.load(f)


def apply(tuning: dict | None = None) -> None:
    """Apply a tuning table (default: tuning.json) to this process."""
    from . import primitives as P
    from ._lib import SORT_KEYS_ONLY, VxError
    t = load() if tuning is None else tuning
    for key, kb in (("sort_u32", 4), ("sort_u64", 8), ("sort_u64_keys", 8 | SORT_KEYS_ONLY)):
        if key in t:
            try:
                P.set_sort_tuning(kb, *t[key])
            except VxError:
                pass  # shape not supported on this device: keep the default
    if "scan" in t:
        try:
            P.set_scan_tuning(*t["scan"])
        except VxError:
            pass


def _timer():
    import torch
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def timed(fn, reps=5):
        fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            e0.record()
            fn()
            e1.record()
            e1.synchronize()
            ts.append(e0.elapsed_time(e1))
        ts.sort()
        return ts[len(ts) // 2]

    return timed


def sweep_sort(log2_keys: int, key_bytes: int, key_bits: int, with_values: bool = True) -> list:
    import torch

    from . import primitives as P
    from ._lib import SORT_KEYS_ONLY, VxError
    n = 1 << log2_keys
    timed = _timer()
    if key_bytes == 4:
        src_k = P.gen_u32(2, 1, n, 0, 0xFFFFFFFF)
    elif key_bits == 42:
        src_k = P.gen_morton(2, 1, n, 14)
    else:
        src_k = P.gen_morton(2, 1, n, key_bits // 3)
    src_v = P.iota(n) if with_values else None
    k = torch.empty_like(src_k)
    v = torch.empty_like(src_v) if with_values else None
    ws = torch.empty(P.required_sort_workspace_size(n, key_bytes, key_bits, True) * 2 + (1 << 20),
                     dtype=torch.uint8, device="cuda")
    row = key_bytes if with_values else key_bytes | SORT_KEYS_ONLY

    def copy_in():
        k.copy_(src_k)
        if with_values:
            v.copy_(src_v)

    base = timed(copy_in)
    out = []
    for rb, kpt, blk in itertools.product(RADIX_BITS, KEYS_PER_THREAD, BLOCKS):
        try:
            P.set_sort_tuning(row, rb, kpt, blk)
        except VxError:
            continue

        def run():
            copy_in()
            P.sort_by_key(k, v, key_bits, workspace=ws)

        ms = timed(run) - base
        out.append({"radix_bits": rb, "keys_per_thread": kpt, "block": blk, "ms": ms,
                    "keys_per_s": n / ms * 1e3})
    return out


def sweep_scan(log2_n: int) -> list:
    import torch

    from . import primitives as P
    from ._lib import VxError
    n = 1 << log2_n
    timed = _timer()
    src = P.gen_u32(2, 3, n, 0, 15)
    a = torch.empty_like(src)
    ws = torch.empty(max(1, P.required_workspace_size(n)), dtype=torch.uint8, device="cuda")
    base = timed(lambda: a.copy_(src))
    out = []
    for blk, kv in SCAN_SHAPES:
        try:
            P.set_scan_tuning(blk, kv)
        except VxError:
            continue
        ms = timed(lambda: (a.copy_(src), P.exclusive_sum(a, workspace=ws))) - base
        out.append({"block": blk, "vectors_per_lane": kv, "ms": ms, "gbs": 8 * n / ms / 1e6})
    return out


def sweep_batched(workload) -> list:
    """Sort shapes on the batched edit: workload() returns (Batcher, rounds)."""
    import time

    import torch

    from . import primitives as P
    from ._lib import VxError
    out = []
    for rb, kpt, blk in itertools.product(RADIX_BITS, KEYS_PER_THREAD, BLOCKS):
        try:
            P.set_sort_tuning(8, rb, kpt, blk)
        except VxError:
            continue
        ex, rounds = workload()
        ex.run(rounds[0])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for r in rounds[1:]:
            ex.run(r)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / max(1, len(rounds) - 1)
        out.append({"radix_bits": rb, "keys_per_thread": kpt, "block": blk, "ms": ms})
        ex.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--sort-keys", type=int, default=26, help="log2 of the key count of the sort sweep")
    ap.add_argument("--scan-elems", type=int, default=28)
    ap.add_argument("--record-keys", type=int, default=24,
                    help="log2 of the key count of the packed record-key sweep (an edit of 4096 brushes at "
                         "16384^3 sorts about 2^24 records)")
    ap.add_argument("--out", default=TUNING_PATH)
    args = ap.parse_args()
    sys.path.insert(0, os.path.dirname(HERE))
    rows = ("sort_u32", "sort_u64", "sort_u64_keys")
    res = {"sort_u32": sweep_sort(args.sort_keys, 4, 32), "sort_u64": sweep_sort(args.sort_keys - 1, 8, 42),
           "sort_u64_keys": sweep_sort(args.record_keys, 8, 36, with_values=False),
           "scan": sweep_scan(args.scan_elems)}
    best = {k: min(res[k], key=lambda r: r["ms"]) for k in res}
    table = {k: [best[k]["radix_bits"], best[k]["keys_per_thread"], best[k]["block"]] for k in rows}
    table["scan"] = [best["scan"]["block"], best["scan"]["vectors_per_lane"]]
    table["sweep"] = res
    with open(args.out, "w") as f:
        json.dump(table, f, indent=1)
    print(json.dumps({k: table[k] for k in rows + ("scan",)}))
    for k in rows + ("scan",):
        for r in sorted(res[k], key=lambda r: r["ms"]):
            print(k, r)


if __name__ == "__main__":
    main()
