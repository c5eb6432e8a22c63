#!/usr/bin/env python3
"""MSM autotuner, one size class at a time: the Pippenger window width
(--what window) or the accumulation's slices per resident thread, the bucket
tile (--what rounds).

For each log2(n) in the range, commits `columns` witness-like columns of n
scalars (the shards' scalar mix) resident in HBM at each candidate value and
keeps the fastest. Every value gives the same commitments (checked), so the
choice is purely a speed knob; the library's built-in tables
(msm.hip: msm_pick_c, msm_pick_rounds) hold the results of these sweeps on
MI355X.

    python apogee-vm_amd/autotune.py [--what window|rounds] [--lo 8] [--hi 22] [--columns 8] [--reps 5]

Prints one JSON object {log2n: {"best": value, "ms": {value: ms, ...}}} on stdout.
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import sys
import time

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def tune_class(apg, ctx, lg: int, columns: int, reps: int, candidates, seed: int = 77, what: str = "window"):
    n = 1 << lg
    tau = apg.synthetic(seed, 1, "uniform")[0]
    srs = apg.Srs.generate(tau, n, ctx)
    buf = apg.DeviceBuffer(ctx, 32 * n * columns)
    for c in range(columns):
        apg.synthetic_device(ctx, seed + 1000 + c, n, "witness", out=buf, offset_elems=c * n)
    ctx.synchronize()
    times, ref = {}, None
    try:
        for c in candidates:
            if what == "window":
                apg.set_window_bits(lg, c)
            else:
                apg.set_slice_rounds(lg, c)
            srs.prepare(n)
            out = apg.commit_device(srs, buf, n, columns, ctx)  # warm-up (and table build)
            if ref is None:
                ref = out
            elif out != ref:
                raise RuntimeError(f"{what} {c} changed the commitments at n = 2^{lg}")
            t0 = time.perf_counter()
            for _ in range(reps):
                apg.commit_device(srs, buf, n, columns, ctx)
            times[c] = 1e3 * (time.perf_counter() - t0) / reps
            print(f"[autotune] 2^{lg} x {columns}: {what} = {c}: {times[c]:.3f} ms", file=sys.stderr, flush=True)
    finally:
        apg.set_window_bits(lg, 0)
        apg.set_slice_rounds(lg, 0)
        srs.release()
        buf.free()
    best = min(times, key=times.get)
    return best, times


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--lo", type=int, default=8)
    ap.add_argument("--hi", type=int, default=22)
    ap.add_argument("--columns", type=int, default=8)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--span", type=int, default=3, help="window candidates: default c +- span")
    ap.add_argument("--what", default="window", choices=["window", "rounds"])
    args = ap.parse_args()
    apg = importlib.import_module("apogee-vm_amd")
    ctx = apg.Context(0)
    table = {}
    for lg in range(args.lo, args.hi + 1):
        if args.what == "window":
            base = apg.window_bits(1 << lg)
            cands = [c for c in range(base - args.span, base + args.span + 1) if 4 <= c <= 17]
        else:
            base = apg.slice_rounds(1 << lg)
            cands = [2, 3, 4, 6, 8]
        best, times = tune_class(apg, ctx, lg, args.columns, args.reps, cands, what=args.what)
        table[lg] = {"best": best, "default": base, "ms": {str(c): round(t, 4) for c, t in times.items()}}
    print(json.dumps(table), flush=True)


if __name__ == "__main__":
    main()
