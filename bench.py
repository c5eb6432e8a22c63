#!/usr/bin/env python3
"""Benchmark: per-shard commitment MSM + sumcheck on one MI355X.

Workloads (--workload):
  mixed  (default) the metric's batch: 207 shards from 23 families of 4-24
         witness-like columns, log2(rows) 60% in 8-14, 30% in 15-19, 10% in
         20-22, each shard = its column commitments + a degree-3 batched
         sumcheck (apg_prove_shard_dev). A step is one pass over the batch
         through the shard engine (concurrent workers, continuous batching);
         the plain one-shard-per-call path is timed on the same inputs and
         every shard's bytes are checked equal between the two.
  shard  one shard of 2^20 rows x 8 columns (the single large shard), a step
         per shard.
  serve  the batch's shards arriving as a seeded Poisson process at 0.5x,
         1x and 1.5x the measured batched capacity: shards/s, p50/p99/p999
         latency (and latency over the shard's isolated time), HBM occupancy.
  copy   1 GiB device copies by the library's k_copy16 kernel: the
         known-size reference for rocprofv3 FETCH_SIZE / WRITE_SIZE passes.
  open   pcs::open (the Mercury opening) of one witness-mix column of
         2^--open-vars rows resident in HBM, a step per opening; the oracle
         opens a 2^16 column on the host beside it (parity checked).
  gkr    one GKR layer transition at 2^--gkr-vars rows in HBM (16 columns,
         20 weighted gates of the reference's shapes, two virtual tables):
         the forward pass, the row-wise layer sumcheck and the halving list's
         sumcheck over the next layer; the oracle runs the row-wise sumcheck
         at 2^14 rows beside it (parity checked).
  srs    Srs::load of a 2^--srs-power archive (default 2^22) with every G1
         point validated on the device, and the decode of the same block from
         host memory; the oracle's decoder runs on a 2^14-point sample.
Inputs are generated into HBM before any timed region. One JSON line on
stdout; progress on stderr. The mixed run reports the bucket-accumulation
kernel's roofline over the engine's timed passes (HIP events on each launch's
stream; also its average over every launch of the process, and alone in a
single 2^20 x 8 shard), the single large shard's numbers, and a CPU baseline:
the oracle's restatement of the reference algorithms timed on a sample of the
batch's shards on the host cores, whose outputs are checked against the GPU's.

    python bench.py [--workload mixed|shard|serve|copy|open|gkr|srs] [--steps K] [--warmup W]
                    [--workers N] [--policy fifo|shortest|longest] [--reserve R] [--lane-cus C]
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# One hardware queue per engine worker stream (plus this process's context):
# with HIP's default of 4, streams share in-order queues and a short shard's
# kernels wait behind a long shard's. Read once, when HIP starts.
os.environ["GPU_MAX_HW_QUEUES"] = "16"

METRIC = "shards/s (commit MSM+sumcheck) on mixed 2^8–2^22 batch; G1 MSM pts/s; p99 lat"
# v_mad_u64_u32 (32x32->64 multiply-add) issues at the FP64-FMA rate on gfx950:
# 64 lanes per CU per clock, half the rate of 32-bit adds and logic. At
# 2.4 GHz that is 256 x 64 x 2.4e9 = 3.93e13/s, the published FP64 vector
# 78.6 TFLOP/s counted as FMAs; 32 independent chains per lane sustain
# 3.3e13/s under load.
PEAK_IMUL = 256 * 64 * 2.4e9
PEAK_HBM = 8.0e12
IMUL_PER_MADD = 1496  # madd-2007-bl: 7M + 4S Fq multiplies x 136
# serving engine: workers (one of them the fast lane) and the preemption ratio
SERVE_WORKERS = 6
SERVE_RESERVE = 2
SERVE_LANE_CUS = 0
SERVE_PREEMPT = 2.0
SERVE_PREEMPT_MODE = "size"
SERVE_URGENCY_RATIO = 1.2

apg = None
wl = None


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def fill_shard(ctx, spec):
    buf = apg.DeviceBuffer(ctx, 32 * spec.elements)
    for c in range(spec.columns):
        apg.synthetic_device(ctx, spec.seed + c, spec.rows, "witness", out=buf, offset_elems=c * spec.rows)
    return buf


def same(p, q):
    return p.commitments == q.commitments and np.array_equal(p.rounds, q.rounds) and np.array_equal(
        p.final_evals, q.final_evals)


# ------------------------------------------------------------------ single large shard
def run_single(ctx, srs, args, nv=20, k=8, deg=3, steps=10, warmup=2):
    n = 1 << nv
    spec = wl.ShardSpec(0, 0, nv, k, deg, args.seed + 77)
    buf = fill_shard(ctx, spec)
    for _ in range(warmup):
        apg.prove_shard(srs, buf, nv, k, deg, apg.Transcript(), ctx)
    ctx.synchronize()
    # timed steps without kernel events (they cost microseconds per launch);
    # the library's host clocks give the MSM and sumcheck shares
    was = ctx.profiling
    ctx.set_profiling(False)
    before = ctx.stats()
    lat = []
    t0 = time.perf_counter()
    for _ in range(steps):
        ts = time.perf_counter()
        proof = apg.prove_shard(srs, buf, nv, k, deg, apg.Transcript(), ctx)
        lat.append(time.perf_counter() - ts)
    ctx.synchronize()
    total = time.perf_counter() - t0
    wall = stats_delta(ctx.stats(), before)
    # then the same steps with kernel events: the kernels' own times
    ctx.set_profiling(True)
    before = ctx.stats()
    for _ in range(steps):
        apg.prove_shard(srs, buf, nv, k, deg, apg.Transcript(), ctx)
    ctx.synchronize()
    st = stats_delta(ctx.stats(), before)
    ctx.set_profiling(was)
    roof = roofline(st, f"k_accumulate (MSM bucket accumulation), one 2^{nv} x {k} shard alone")
    sc_gbs = st["sumcheck_bytes"] / (st["sumcheck_kernel_ms"] * 1e-3) / 1e9 if st["sumcheck_kernel_ms"] else 0.0
    # the survey's algorithmic multiplies of a sumcheck over eq and deg - 1
    # tables (t = deg live tables): 136 ((t + 1)(t - 1) + t) 2^n IMUL
    live = deg
    sc_imul = 136 * ((live + 1) * (live - 1) + live) * n
    sc_imul_rate = sc_imul * steps / (st["sumcheck_kernel_ms"] * 1e-3) if st["sumcheck_kernel_ms"] else 0.0
    # and its bytes, 96 t 2^n + 32 2^n (eq counted as a table; the library
    # never materialises eq, so its own count above is lower)
    sc_survey_bytes = (96 * live + 32) * n
    sc_survey_gbs = sc_survey_bytes * steps / (st["sumcheck_kernel_ms"] * 1e-3) / 1e9 if st["sumcheck_kernel_ms"] else 0.0
    ff = fixed_formula_msm(buf, n, k)
    msm_s = wall["msm_ms"] * 1e-3 / steps
    ff_rate = ff["imul"] / msm_s if msm_s else 0.0
    res = {
        "workload": f"one shard: 2^{nv} rows x {k} witness columns, degree-{deg} sumcheck",
        "shards_per_s": round(steps / total, 3),
        "ms_per_shard": round(1e3 * total / steps, 3),
        "g1_msm_points_per_s": round(wall["msm_points"] / (wall["msm_ms"] * 1e-3), 1) if wall["msm_ms"] else None,
        "msm_ms_per_shard": round(wall["msm_ms"] / steps, 3),
        "sumcheck_ms_per_shard": round(wall["sumcheck_ms"] / steps, 3),
        "host_poseidon2_us": host_permutation_us(),
        "sumcheck_kernel_gbs": round(sc_gbs, 1),
        "sumcheck_hbm_frac": round(sc_gbs * 1e9 / PEAK_HBM, 4),
        "sumcheck_kernel_gbs_survey_bytes": round(sc_survey_gbs, 1),
        "sumcheck_kernel_timul_per_s": round(sc_imul_rate / 1e12, 3),
        "sumcheck_imul_frac": round(sc_imul_rate / PEAK_IMUL, 4),
        "latency_ms": {"p50": round(float(np.percentile(np.array(lat) * 1e3, 50)), 3),
                       "p99": round(float(np.percentile(np.array(lat) * 1e3, 99)), 3)},
        # the whole MSM (every kernel and the host tail, wall clock) against
        # the survey's fixed algorithmic work at the reference's window width
        "msm_fixed_formula": {**ff, "achieved_timul_per_s": round(ff_rate / 1e12, 3),
                              "frac_of_peak": round(ff_rate / PEAK_IMUL, 4)},
    }
    return res, roof, (spec, buf, proof)


def host_permutation_us(n=2000):
    """Microseconds per Poseidon2 permutation on this host, the transcript's
    dispatched multiply (form 0) and the portable one (form 1): every round of
    a sumcheck waits on a few of these."""
    st = np.zeros((n, 3, 4), dtype=np.uint64)
    st[:, :, 0] = np.arange(3 * n, dtype=np.uint64).reshape(n, 3)
    out = np.zeros_like(st)
    res = {}
    for form, name in ((0, "dispatched"), (1, "portable")):
        best = float("inf")
        for _ in range(5):
            t0 = time.perf_counter()
            rc = apg.lib().apg_debug_poseidon2_batch(form, st.ctypes.data, n, out.ctypes.data)
            best = min(best, time.perf_counter() - t0)
            assert rc == 0
        res[name] = round(best / n * 1e6, 3)
    return res


def stats_delta(after, before):
    return {k: after[k] - before[k] for k in after}


def stats_sum(a, b):
    return {k: a[k] + b[k] for k in a}


def traffic_per_entry():
    """HBM bytes per accumulated digit entry of k_accumulate, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE passes (scaled by the same
    counters on the k_copy16 kernel); None when no such summary exists."""
    path = os.path.join(ROOT, "profiles", "k_accumulate_traffic.md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("bytes_per_entry:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def roofline(st, label):
    """Roofline of k_accumulate over the launches recorded in `st`: the
    algorithmic multiplies (1496 per digit entry: one mixed addition) per
    launch over the average launch time (HIP events on the launch stream)."""
    launches = max(1, st["accumulate_launches"])
    avg_ms = st["accumulate_ms"] / launches
    entries = st["accumulate_entries"] / launches
    achieved = IMUL_PER_MADD * entries / (avg_ms * 1e-3) if avg_ms else 0.0
    tpe = traffic_per_entry()
    return {
        "kernel": label,
        "bound": "valu-imul",
        "achieved": round(achieved / 1e12, 3),
        "peak": round(PEAK_IMUL / 1e12, 3),
        "unit": "TIMUL/s",
        "frac": round(achieved / PEAK_IMUL, 4),
        "traffic": round(tpe * entries) if tpe is not None else None,
        "launches": int(st["accumulate_launches"]),
        "avg_launch_ms": round(avg_ms, 4),
        "entries_per_launch": round(entries, 1),
        "algorithmic_imul_per_launch": round(IMUL_PER_MADD * entries),
    }


# The survey's fixed algorithmic MSM work (independent of this library's
# window choice): per column, 1496 IMUL per nonzero signed digit under the
# reference's rule w = floor(0.69 log2 n) + 2 (crates/curve/src/msm.rs:112-118,
# 180-198), plus ceil(255 / w) windows of bucket reduction (2 * 2^(w-1)
# general additions of 2176 IMUL) and w doublings (952 IMUL).
FR_P = 0x30644E72E131A029B85045B68181585D2833E84879B9709143E1F593F0000001
R_INV = pow(2**256, -1, FR_P)


def ref_window(n):
    if n < 32:
        return 3
    lg = n.bit_length() - 1 if n & (n - 1) == 0 else n.bit_length()
    return lg * 69 // 100 + 2


def nonzero_digits(scalars_mont, w):
    """Nonzero signed digits of Montgomery Fr scalars ((m, 4) uint64) under
    the reference's recoding at width w."""
    canon = [int.from_bytes(row.tobytes(), "little") * R_INV % FR_P for row in scalars_mont]
    windows = (254 + 1 + w - 1) // w
    half, mask = 1 << (w - 1), (1 << w) - 1
    v = np.array([[(x >> (w * j)) & mask for j in range(windows)] for x in canon], dtype=np.int64)
    carry = np.zeros(v.shape[0], dtype=np.int64)
    nz = 0
    for j in range(windows):
        coef = v[:, j] + carry
        if j == windows - 1:
            nz += int(np.count_nonzero(coef))
        else:
            big = coef >= half
            nz += int(np.count_nonzero(np.where(big, coef - (1 << w), coef)))
            carry = big.astype(np.int64)
    return nz


def fixed_formula_msm(buf, n, k, sample=1 << 15):
    """The survey's algorithmic IMUL of k column MSMs of n points, the digit
    count estimated from the first `sample` scalars of every column."""
    w = ref_window(n)
    host = buf.download().reshape(k, n, 4)
    m = min(sample, n)
    nz = sum(nonzero_digits(host[c, :m], w) for c in range(k)) * (n / m)
    windows = -(-255 // w)
    imul = IMUL_PER_MADD * nz + k * windows * (2 * 2 ** (w - 1) * 2176 + w * 952)
    return {"w_ref": w, "nonzero_digits_per_point": round(nz / (n * k), 3), "imul": float(imul),
            "sample": f"digits counted on the first {m} scalars of each column, scaled by n / {m}"}


# ------------------------------------------------------------------ mixed batch
def setup_batch(ctx, args):
    specs = wl.mixed_batch(seed=args.seed)
    nmax = max(s.rows for s in specs)
    tau = apg.synthetic(args.seed, 1, "uniform")[0]
    t0 = time.perf_counter()
    srs = apg.Srs.generate(tau, max(nmax, 1 << 20), ctx)
    for nv in sorted({s.num_vars for s in specs} | {20}, reverse=True):  # largest first: one table per width
        srs.prepare(1 << nv)
    log(f"SRS 2^{max(nmax, 1 << 20).bit_length() - 1} generated, window tables for "
        f"{len({s.num_vars for s in specs})} size classes in {time.perf_counter() - t0:.2f} s; "
        f"{srs.device_bytes() / 2**30:.2f} GiB resident")
    t0 = time.perf_counter()
    bufs = [fill_shard(ctx, s) for s in specs]
    ctx.synchronize()
    tot = sum(s.elements for s in specs)
    log(f"batch: {len(specs)} shards, {tot / 1e6:.1f} M scalars ({32 * tot / 2**30:.2f} GiB) in HBM "
        f"in {time.perf_counter() - t0:.2f} s")
    return specs, srs, bufs


def run_plain(ctx, srs, specs, bufs):
    proofs, iso = [], []
    t0 = time.perf_counter()
    for s, b in zip(specs, bufs):
        ts = time.perf_counter()
        proofs.append(apg.prove_shard(srs, b, s.num_vars, s.columns, s.degree, apg.Transcript(), ctx))
        iso.append(time.perf_counter() - ts)
    ctx.synchronize()
    return time.perf_counter() - t0, proofs, np.array(iso)


def run_engine_pass(eng, srs, specs, bufs):
    t0 = time.perf_counter()
    ids = [eng.submit_shard(srs, b, s.num_vars, s.columns, s.degree) for s, b in zip(specs, bufs)]
    out = [eng.wait(j) for j in ids]
    return time.perf_counter() - t0, out


def host_threads():
    """The host threads this job may use: the CPUs it is allowed to run on, at
    most 16 (a one-GPU machine's share of its host's cores)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(srs, specs, bufs, gpu_proofs, budget_s=25.0):
    """The oracle (the restated reference algorithms) on a sample of the
    batch, every shard's outputs compared with the GPU's: the narrowest shard
    of 2^22 and of 2^20 rows first, then shards of at most 2^18 rows in a
    fixed shuffled order until about budget_s of CPU work. Scaled to the whole
    batch per size class: a shard costs its elements times the sampled
    seconds per element of its log2(rows), interpolated in log2(rows) between
    sampled sizes (the per-element cost of Pippenger falls with n)."""
    import oracle  # the checker, used only for this baseline leg

    cores = host_threads()
    big = []
    for nv in (22, 20):
        cand = [i for i, s in enumerate(specs) if s.num_vars == nv]
        if cand:
            big.append(min(cand, key=lambda i: specs[i].columns))
    rng = np.random.default_rng(5)
    small = [int(i) for i in rng.permutation(len(specs)) if specs[i].num_vars <= 18]
    nmax = max(specs[i].rows for i in big) if big else 1 << 18
    srs_aff = srs.g1(0, nmax)
    used, done, ok = 0.0, [], True
    per_class = {}  # log2 rows -> [seconds, elements]
    for i in big + small:
        s = specs[i]
        cols = bufs[i].download().reshape(s.columns, s.rows, 4)
        t = time.perf_counter()
        cms, rounds, fin = oracle.shard_script(srs_aff[:s.rows], cols, s.degree, oracle.Transcript(), workers=cores)
        dt = time.perf_counter() - t
        used += dt
        acc = per_class.setdefault(s.num_vars, [0.0, 0])
        acc[0] += dt
        acc[1] += s.elements
        done.append(int(i))
        g = gpu_proofs[i]
        ok = ok and g.commitments == cms and np.array_equal(g.rounds, rounds) and np.array_equal(g.final_evals, fin)
        if used > budget_s and len(done) > len(big):
            break
    xs = np.array(sorted(per_class))
    ys = np.log(np.array([per_class[x][0] / per_class[x][1] for x in xs]))
    est_total = sum(s.elements * float(np.exp(np.interp(s.num_vars, xs, ys))) for s in specs)
    return {
        "value": round(len(specs) / est_total, 4),
        "unit": "shards/s",
        "cores": cores,
        "kind": "port",
        "sample": f"{len(done)} of the batch's shards (one of 2^22 and one of 2^20 rows, the rest of at most "
                  f"2^18) in {used:.1f} s on {cores} threads, scaled to all {len(specs)} by seconds per column "
                  f"element per size class (log2 rows {int(xs.min())}-{int(xs.max())} sampled, interpolated)",
    }, ok, len(done)


def run_peak(ctx, args):
    """The measured issue rate of v_mad_u64_u32 against the published peak
    the rooflines divide by."""
    r = apg.imul_peak(ctx, iters=4096, reps=max(3, args.steps))
    return {"metric": "v_mad_u64_u32 issue rate (32 independent chains per lane, 8 blocks of 256 per CU)",
            "value": round(r["imul_per_s"] / 1e12, 3), "unit": "TIMUL/s", "n_gpus": 1, "steps": max(3, args.steps),
            "warmup": 1, "ms_per_step": round(r["best_ms"], 4), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "u32 x u32 -> u64", "data": "synthetic",
            "config": {"workload": "issue-rate probe", "parallelism": "1 GPU"},
            "published_peak_timul_per_s": round(PEAK_IMUL / 1e12, 3),
            "frac_of_published": round(r["imul_per_s"] / PEAK_IMUL, 4)}


def msm_uniform_line(ctx, srs, seed, reps=5):
    """G1 MSM points/s of one 2^22-point MSM over uniform Fr scalars, the
    shape of the reference's published figure (pcs::commit at 2^22, 1.30 s on
    18 CPU cores: 3.23 M points/s)."""
    n = 1 << 22
    buf = apg.synthetic_device(ctx, seed + 4242, n, "uniform")
    apg.commit_device(srs, buf, n, 1, ctx)
    ctx.synchronize()
    t = []
    for _ in range(reps):
        t0 = time.perf_counter()
        apg.commit_device(srs, buf, n, 1, ctx)
        t.append(time.perf_counter() - t0)
    buf.free()
    best = min(t)
    return {"workload": "one MSM of 2^22 points, uniform Fr scalars", "ms": round(best * 1e3, 3),
            "g1_points_per_s": round(n / best, 1), "reference_published_points_per_s": 3.23e6,
            "reference_published": "1.30 s for pcs::commit at 2^22 on an 18-core CPU (docs/spec/mercury.md:290)"}


def copy_calibration(ctx, args, nbytes=1 << 30):
    """The known-size reference for the HBM traffic counters: k_copy16 reads
    and writes nbytes per launch; a FETCH_SIZE / WRITE_SIZE pass over this
    run gives the counter's reading per byte moved."""
    src = apg.DeviceBuffer(ctx, nbytes)
    dst = apg.DeviceBuffer(ctx, nbytes)
    apg.synthetic_device(ctx, args.seed, nbytes // 32, "uniform", out=src)
    for _ in range(max(1, args.warmup)):
        dst.copy_from(src)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dst.copy_from(src)
    dt = (time.perf_counter() - t0) / args.steps
    return {"metric": "k_copy16 device copy", "value": round(2 * nbytes / dt / 1e9, 1), "unit": "GB/s",
            "bytes_read_per_launch": nbytes, "bytes_written_per_launch": nbytes, "steps": args.steps,
            "ms_per_step": round(1e3 * dt, 4), "frac_of_hbm_peak": round(2 * nbytes / dt / PEAK_HBM, 4)}


def run_open(ctx, args):
    """pcs::open of one witness-mix column of 2^open_vars rows resident in
    HBM: warmup, then `steps` timed openings (a fresh transcript each), every
    proof compared with the first. The CPU leg is the oracle's opening of a
    2^16 column (the same inputs opened on the GPU too, parity checked),
    scaled to the timed size by rows."""
    nv = args.open_vars
    n = 1 << nv
    tau = apg.synthetic(args.seed, 1, "uniform")[0]
    srs = apg.Srs.generate(tau, n, ctx)
    f = apg.synthetic_device(ctx, args.seed + 1, n, "witness")
    cm = apg.g1_to_wire(apg.commit_device(srs, f, n, 1, ctx)[0])
    u = apg.synthetic(args.seed + 2, nv, "uniform")
    log(f"open: 2^{nv} rows, warmup {args.warmup}")
    first = None
    for _ in range(max(args.warmup, 1)):
        _, first = apg.open(srs, f, cm, u, apg.Transcript(), ctx, num_vars=nv)
    ctx.synchronize()
    t0 = time.perf_counter()
    same_all = True
    for _ in range(args.steps):
        _, proof = apg.open(srs, f, cm, u, apg.Transcript(), ctx, num_vars=nv)
        same_all = same_all and proof == first
    ctx.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    log(f"open: {dt * 1e3:.2f} ms per opening")
    cpu = None
    if args.cpu_budget > 0:
        import oracle
        snv = 16
        sf = apg.synthetic(args.seed + 3, 1 << snv, "witness")
        su = apg.synthetic(args.seed + 4, snv, "uniform")
        ssrs = apg.Srs.generate(tau, 1 << snv, ctx)
        powers = ssrs.g1()
        scm = apg.g1_to_wire(apg.commit(ssrs, sf))
        _, gpu_proof = apg.open(ssrs, sf, scm, su, apg.Transcript(), ctx)
        workers = host_threads()
        t1 = time.perf_counter()
        _, cpu_proof = oracle.mercury_open(powers, sf, scm, su, oracle.Transcript(), workers=workers)
        cpu_s = time.perf_counter() - t1
        cpu = {"value": round((1 << snv) / n / cpu_s, 5), "unit": "openings/s", "cores": workers, "kind": "port",
               "sample": f"one 2^{snv}-row opening in {cpu_s:.2f} s, scaled to 2^{nv} rows by row count",
               "parity": gpu_proof == cpu_proof}
    return {"metric": f"Mercury openings/s (pcs::open of one 2^{nv}-row column resident in HBM)",
            "value": round(1.0 / dt, 3), "unit": "openings/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 3), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "u32x8 (BN254 Fr/Fq Montgomery, exact)",
            "data": "synthetic (counter-mode ChaCha20 witness mix in HBM)",
            "config": {"workload": f"pcs::open, one column of 2^{nv} rows", "parallelism": "1 GPU"},
            "deterministic": same_all, "cpu_baseline": cpu}


# ------------------------------------------------------------------ GKR layer
GKR_COLS = 16
FR_MUL_IMUL = 136  # one 8 x 32-bit CIOS Montgomery multiply: 64 + 64 + 8


def gkr_artifact(g, ncols=GKR_COLS):
    """A two-list circuit of the reference's shapes over `ncols` committed
    columns: a row-wise list (8 products, 4 affine products of 3 + 3 terms,
    2 quadratics, 2 linear copies, 4 enforcing quadratics, the RowIndex and
    Range16 virtual tables read by two of them), then a halving list (a
    product and a fraction tree over its output: tree products and crosses)."""
    lit = lambda x: g.Literal.of(apg.fr_from_u64(x))
    ch = g.Challenge(1)
    W = [g.Witness(i) for i in range(ncols)]
    rowi, r16 = g.Virtual(g.VirtualKind.RowIndex), g.Virtual(g.VirtualKind.Range16)
    prod = []
    for j in range(8):
        prod.append(g.Product(lit(j + 2), W[j], W[(j + 5) % ncols]))
    for j in range(4):
        a, b = [W[(3 * j + q) % ncols] for q in range(3)], [W[(3 * j + 7 + q) % ncols] for q in range(3)]
        prod.append(g.AffineProduct([(lit(q + 2), x) for q, x in enumerate(a)], lit(3),
                                    [(ch if q == 0 else lit(q + 5), x) for q, x in enumerate(b)], lit(7)))
    prod.append(g.Quadratic(lit(7), [(lit(2), W[1]), (ch, rowi)],
                            [(lit(3), W[2], W[3]), (ch, W[4], W[5]), (lit(5), W[6], r16)]))
    prod.append(g.Quadratic(lit(1), [(lit(9), W[8]), (lit(4), W[9])],
                            [(lit(3), W[10], W[11]), (lit(6), W[12], W[13]), (lit(5), W[14], W[15])]))
    prod.append(g.Linear([(lit(1), W[14])], lit(0)))
    prod.append(g.Linear([(lit(1), W[15]), (lit(2), W[0])], lit(0)))
    enf = [g.Quadratic(lit(j + 1), [(lit(2), W[j])], [(lit(3), W[j + 4], W[j + 8])]) for j in range(4)]
    L1 = [g.Inner(1, j) for j in range(len(prod))]
    row = g.LayerSpec(False, 0, len(prod), producing=[(f"p{j}", L1[j], gt) for j, gt in enumerate(prod)],
                      enforcing=[(f"e{j}", gt) for j, gt in enumerate(enf)])
    half = []
    for j in range(0, len(prod), 2):
        half.append(g.TreeProduct(L1[j]))
        half.append(g.TreeCross(L1[j + 1], L1[j]))
    halving = g.LayerSpec(True, 0, len(half),
                          producing=[(f"h{j}", g.Inner(2, j), gt) for j, gt in enumerate(half)])
    art = g.CircuitArtifact([], [f"w{i}" for i in range(ncols)], [], [(g.VirtualKind.RowIndex, "row"),
                                                                     (g.VirtualKind.Range16, "r16")],
                            [row, halving])
    return art, {1: apg.synthetic(99, 1, "uniform")[0]}


def gkr_mults_per_pair(res):
    """Fr multiplies per pair of one round (three nodes; the weight and eq
    products; the fold of every table), from the flattened list."""
    per_node = 1  # E(i) S
    for shape, s0, s1, _, t in res.gates:
        nv = int(s1 - s0)
        m = {0: nv, 1: 2, 2: 1, 3: nv + 1, 4: 1, 5: int(t) + 2 * ((nv - int(t)) // 2), 6: 2}[int(shape)]
        per_node += m
    per_node += res.nproducing + res.nenforcing  # weights
    return 3 * per_node + 1  # + the eq factor


class _View:
    def __init__(self, buf, offset_bytes):
        import ctypes
        self.ptr = ctypes.c_void_p(buf.ptr.value + offset_bytes)


def run_gkr(ctx, args):
    """One GKR layer transition at 2^gkr_vars rows in HBM: the forward pass of
    the row-wise list (layer 0 -> layer 1), its layer sumcheck, and the
    halving list's sumcheck over layer 1's children. A step is the three. The
    CPU leg runs the oracle's row-wise sumcheck at 2^14 rows (parity checked
    against the GPU there) and is scaled to the timed size by rows."""
    g = importlib.import_module("apogee-vm_amd.gkr")
    nv = args.gkr_vars
    n = 1 << nv
    art, chal = gkr_artifact(g)
    res, resh = g.ResolvedList(art, 0, chal), g.ResolvedList(art, 1, chal)
    cols = [apg.synthetic_device(ctx, args.seed + c, n, "witness") for c in range(GKR_COLS)]
    outs = [apg.DeviceBuffer(ctx, 32 * n) for _ in range(res.nproducing)]
    w = apg.synthetic(args.seed + 100, res.nproducing + res.nenforcing, "uniform")
    wh = apg.synthetic(args.seed + 101, resh.nproducing, "uniform")
    r = apg.synthetic(args.seed + 102, nv, "uniform")
    rh = apg.synthetic(args.seed + 103, nv - 1, "uniform")
    lower = [_View(b, 0) for b in outs]
    upper = [_View(b, 32 * (n // 2)) for b in outs]

    def step():
        t0 = time.perf_counter()
        g.forward_layer_device(res, cols, nv, outs, ctx)
        t1 = time.perf_counter()
        p1 = g.prove_sumcheck_device(r, res, w, cols, None, apg.Transcript(), ctx)
        t2 = time.perf_counter()
        p2 = g.prove_sumcheck_device(rh, resh, wh, lower, upper, apg.Transcript(), ctx)
        t3 = time.perf_counter()
        return (t1 - t0, t2 - t1, t3 - t2), (p1, p2)

    log(f"gkr: 2^{nv} rows, {GKR_COLS} columns, {res.nproducing} + {res.nenforcing} gates, warmup {args.warmup}")
    first = None
    for _ in range(max(args.warmup, 1)):
        _, first = step()
    ctx.synchronize()
    ctx.set_profiling(True)
    before = ctx.stats()
    times = []
    same_all = True
    t0 = time.perf_counter()
    for _ in range(args.steps):
        tt, proofs = step()
        times.append(tt)
        same_all = same_all and all(np.array_equal(a, b) for p, q in zip(proofs, first) for a, b in zip(p, q))
    ctx.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    st = stats_delta(ctx.stats(), before)
    ctx.set_profiling(False)
    fw, sc, hv = (float(np.mean([t[i] for t in times])) for i in range(3))
    log(f"gkr: forward {fw * 1e3:.2f} ms, row-wise sumcheck {sc * 1e3:.2f} ms, halving sumcheck {hv * 1e3:.2f} ms")
    # roofline of the round kernels of the row-wise sumcheck: Fr multiplies x 136 IMUL
    mults = gkr_mults_per_pair(res)
    pairs_row = n - 1  # sum over rounds of the pairs
    multsh = gkr_mults_per_pair(resh)
    pairs_half = n // 2 - 1
    # the fold: 2 multiplies per table per pair from round 1 on
    folds = 2 * GKR_COLS * (n // 2 - 1) + 2 * 2 * GKR_COLS * (n // 4 - 1)
    imul = FR_MUL_IMUL * (mults * pairs_row + multsh * pairs_half + folds) * args.steps
    kern_s = st["sumcheck_kernel_ms"] * 1e-3
    achieved = imul / kern_s if kern_s else 0.0
    roof = {"kernel": "k_gkr_round (layer sumcheck rounds, row-wise and halving lists)", "bound": "valu-imul",
            "achieved": round(achieved / 1e12, 3), "peak": round(PEAK_IMUL / 1e12, 3), "unit": "TIMUL/s",
            "frac": round(achieved / PEAK_IMUL, 4), "traffic": None,
            "launches": int(st["sumcheck_launches"]),
            "avg_launch_ms": round(st["sumcheck_kernel_ms"] / max(1, st["sumcheck_launches"]), 4),
            "fr_mults_per_pair": {"row_wise": mults, "halving": multsh},
            "table_gbs": round(st["sumcheck_bytes"] / kern_s / 1e9, 1) if kern_s else None}
    cpu = None
    if args.cpu_budget > 0:
        import oracle
        snv = 14
        sres = g.ResolvedList(art, 0, chal)
        scol = [apg.synthetic(args.seed + 500 + c, 1 << snv, "witness") for c in range(GKR_COLS)]
        sr = apg.synthetic(args.seed + 600, snv, "uniform")
        a = sres.arrays()
        flat = oracle.GkrFlat(a["gates"], a["ncached"], a["nproducing"], a["nenforcing"], a["sources"],
                              a["coeffs"], a["kinds"], a["halving"])
        t1 = time.perf_counter()
        rounds_o, point_o, fin_o = oracle.gkr_prove_sumcheck(flat, w, sr, scol, None, oracle.Transcript())
        cpu_s = time.perf_counter() - t1
        rounds_d, point_d, fin_d = g.prove_sumcheck(sr, sres, w, scol, None, apg.Transcript(), ctx)
        parity = bool(np.array_equal(rounds_o, rounds_d) and np.array_equal(point_o, point_d)
                      and np.array_equal(fin_o, fin_d))
        cpu = {"value": round((1 << snv) / cpu_s, 1), "unit": "rows/s", "cores": 1, "kind": "port",
               "sample": f"the row-wise layer sumcheck at 2^{snv} rows in {cpu_s:.2f} s (one thread)",
               "parity": parity}
    return {"metric": f"GKR layer sumcheck rows/s (row-wise list, {GKR_COLS} columns, "
                      f"{res.nproducing} + {res.nenforcing} gates, 2^{nv} rows in HBM)",
            "value": round(n / sc, 1), "unit": "rows/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 3), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "u32x8 (BN254 Fr Montgomery, exact)",
            "data": "synthetic (counter-mode ChaCha20 witness mix in HBM)",
            "config": {"workload": f"GKR layer: forward + row-wise sumcheck + halving sumcheck, 2^{nv} rows",
                       "parallelism": "1 GPU"},
            "forward_ms": round(fw * 1e3, 3), "rowwise_sumcheck_ms": round(sc * 1e3, 3),
            "halving_sumcheck_ms": round(hv * 1e3, 3), "deterministic": same_all, "roofline": roof,
            "cpu_baseline": cpu}


# ------------------------------------------------------------------ SRS ingestion
def run_srs(ctx, args):
    """Srs::load of a 2^srs_power archive (written first from a device-built
    SRS, read back from the page cache) and the canonical-point decode of the
    same block from host memory: points/s, each load's points compared with
    the source. The CPU leg decodes a 2^14-point sample with the oracle's
    validating decoder (one thread), scaled by points."""
    import tempfile
    pw = args.srs_power
    n = 1 << pw
    tau = apg.synthetic(args.seed, 1, "uniform")[0]
    src = apg.Srs.generate(tau, n, ctx)
    g2b = bytes(256)  # infinity: valid, and the G1 block is what is timed
    tmpdir = tempfile.mkdtemp(prefix="apg_srs_")
    path = os.path.join(tmpdir, f"srs-{pw}.srs")
    src.save(path, g2b)
    with open(path, "rb") as f:
        f.seek(280)
        wire = f.read()
    ref = src.g1()
    log(f"srs: 2^{pw} points, archive {os.path.getsize(path) / 2**20:.0f} MiB")
    ok = True
    for _ in range(max(args.warmup, 1)):
        s, _ = apg.Srs.load(path, ctx)
        s.release()
    t_load, t_wire = [], []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        s, _ = apg.Srs.load(path, ctx)
        t_load.append(time.perf_counter() - t0)
        ok = ok and np.array_equal(s.g1(), ref)
        s.release()
        t0 = time.perf_counter()
        s = apg.Srs.from_wire(wire, ctx)
        t_wire.append(time.perf_counter() - t0)
        s.release()
    os.remove(path)
    os.rmdir(tmpdir)
    tl, tw = float(np.median(t_load)), float(np.median(t_wire))
    log(f"srs: load {tl * 1e3:.1f} ms, decode from host memory {tw * 1e3:.1f} ms")
    cpu = None
    if args.cpu_budget > 0:
        import oracle
        m = 1 << 14
        t0 = time.perf_counter()
        dec = [oracle.affine_from_wire(wire[64 * i:64 * i + 64]) for i in range(m)]
        cpu_s = time.perf_counter() - t0
        parity = all(d is not None and np.array_equal(d, ref[i]) for i, d in enumerate(dec))
        cpu = {"value": round(m / cpu_s, 1), "unit": "points/s", "cores": 1, "kind": "port",
               "sample": f"2^14 points through the oracle's validating decoder in {cpu_s:.2f} s (one thread)",
               "parity": parity}
    return {"metric": f"SRS ingestion points/s (Srs::load of a 2^{pw}-point archive, validated on the device)",
            "value": round(n / tl, 1), "unit": "points/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(tl * 1e3, 3), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "u32x8 (BN254 Fq Montgomery, exact)",
            "data": "synthetic (an SRS built on the device, saved and read back)",
            "config": {"workload": f"Srs::load, 2^{pw} G1 points", "parallelism": "1 GPU"},
            "load_gbs": round(64 * n / tl / 1e9, 2), "decode_from_host_ms": round(tw * 1e3, 3),
            "decode_from_host_gbs": round(64 * n / tw / 1e9, 2), "points_equal": ok, "cpu_baseline": cpu}


def main():
    global apg, wl
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="mixed", choices=["mixed", "shard", "serve", "copy", "open", "gkr", "srs", "peak"])
    ap.add_argument("--srs-power", type=int, default=22, help="log2 points of the archive (--workload srs)")
    ap.add_argument("--gkr-vars", type=int, default=20, help="log2 rows of the GKR layer (--workload gkr)")
    ap.add_argument("--open-vars", type=int, default=20, help="log2 rows of the opened table (--workload open)")
    # engine defaults: throughput (mixed) 12 workers in arrival order (8 to 16
    # measured: 12 and more within noise of each other, 1.5% over 8); serving
    # 4 workers (within the hardware queues the device services at once), one
    # of them a fast lane on 32 CUs, shortest job first
    ap.add_argument("--workers", type=int, default=None)
    ap.add_argument("--policy", default=None, choices=["fifo", "shortest", "longest"])
    ap.add_argument("--reserve", type=int, default=None,
                    help="fast-lane workers, kept for shards of at most --small-scalars rows x columns")
    ap.add_argument("--lane-cus", type=int, default=None, help="CUs of the fast lane (0: no CU split)")
    ap.add_argument("--small-scalars", type=float, default=float(1 << 20))
    ap.add_argument("--serve-capacity", type=float, default=0.0,
                    help="shards/s the serving loads are fractions of (0: measure the serving engine's)")
    ap.add_argument("--preempt-mode", default=None, choices=["size", "urgency"],
                    help="what a job gives way to (serving default %s)" % SERVE_PREEMPT_MODE)
    ap.add_argument("--preempt", type=float, default=None,
                    help="phase-boundary preemption ratio (0: off; serving default %g)" % SERVE_PREEMPT)
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of oracle work (0 = skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: multi-GPU runs "
              f"(one process per GPU, each with a contiguous range of SRS points and scalars, "
              f"or whole shards dealt round-robin) are not built", file=sys.stderr)
        sys.exit(2)
    serving = args.workload == "serve"
    for name, mixed_default, serve_default in (("workers", 12, SERVE_WORKERS), ("policy", "fifo", "shortest"),
                                               ("reserve", 0, SERVE_RESERVE), ("lane_cus", 0, SERVE_LANE_CUS),
                                               ("preempt", 0.0, SERVE_PREEMPT),
                                               ("preempt_mode", "size", SERVE_PREEMPT_MODE)):
        if getattr(args, name) is None:
            setattr(args, name, serve_default if serving else mixed_default)
    apg = importlib.import_module("apogee-vm_amd")
    wl = importlib.import_module("apogee-vm_amd.workload")
    ctx = apg.Context(0)

    if args.workload == "copy":
        print(json.dumps(copy_calibration(ctx, args)), flush=True)
        return

    if args.workload == "peak":
        print(json.dumps(run_peak(ctx, args)), flush=True)
        return

    if args.workload == "open":
        print(json.dumps(run_open(ctx, args)), flush=True)
        return

    if args.workload == "gkr":
        print(json.dumps(run_gkr(ctx, args)), flush=True)
        return

    if args.workload == "srs":
        print(json.dumps(run_srs(ctx, args)), flush=True)
        return

    if args.workload == "shard":
        tau = apg.synthetic(args.seed, 1, "uniform")[0]
        srs = apg.Srs.generate(tau, 1 << 20, ctx)
        srs.prepare(1 << 20)
        single, roof, _ = run_single(ctx, srs, args, steps=args.steps, warmup=args.warmup)
        out = {"metric": METRIC, "value": single["shards_per_s"], "unit": "shards/s", "n_gpus": 1,
               "steps": args.steps, "warmup": args.warmup, "ms_per_step": single["ms_per_shard"],
               "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
               "dtype": "u32x8 (BN254 Fr/Fq Montgomery, exact)", "data": "synthetic",
               "config": {"workload": single["workload"], "parallelism": "1 GPU"},
               "single_shard": single, "roofline": roof, "cpu_baseline": None}
        print(json.dumps(out), flush=True)
        return

    specs, srs, bufs = setup_batch(ctx, args)
    # every k_accumulate launch of the process is timed (HIP events on its
    # stream), so the process-wide average compares with a rocprofv3 trace
    ctx.set_profiling(True)
    log("plain path: one shard per call, sequential")
    t_plain, plain, iso = run_plain(ctx, srs, specs, bufs)
    log(f"plain: {len(specs) / t_plain:.1f} shards/s ({t_plain:.3f} s)")
    eng = apg.Engine(0, workers=args.workers, hbm_budget=200 << 30, lane_workers=args.reserve,
                     lane_cus=args.lane_cus, small_scalars=args.small_scalars)
    eng.set_policy(args.policy, args.reserve, args.small_scalars)
    eng.set_preemption(args.preempt, args.preempt_mode)
    eng.set_profiling(True)

    if args.workload == "serve":
        out = serve(eng, srs, specs, bufs, plain, iso, args)
        print(json.dumps(out), flush=True)
        return

    for _ in range(args.warmup):
        run_engine_pass(eng, srs, specs, bufs)
    warm = eng.kernel_stats()
    eng.reset_kernel_stats()
    times, exact = [], True
    for step in range(args.steps):
        dt, res = run_engine_pass(eng, srs, specs, bufs)
        times.append(dt)
        exact = exact and all(same(p, r[0]) for p, r in zip(plain, res))
        log(f"engine pass {step}: {len(specs) / dt:.1f} shards/s ({dt:.3f} s)")
    timed = eng.kernel_stats()
    t_eng = float(np.mean(times))
    info = eng.stats()
    # where the pass's time goes: the engine on each size class alone and on
    # the large and medium shards together, against the plain path's time per
    # class (measured in the plain pass above)
    nvs = np.array([s.num_vars for s in specs])
    classes = {"rows_2e8_2e14": nvs <= 14, "rows_2e15_2e19": (nvs >= 15) & (nvs <= 19), "rows_2e20_2e22": nvs >= 20}
    by_class = {}
    for name, sel in classes.items():
        idx = np.flatnonzero(sel)
        dt, _ = run_engine_pass(eng, srs, [specs[i] for i in idx], [bufs[i] for i in idx])
        by_class[name] = {"shards": int(len(idx)), "engine_ms": round(1e3 * dt, 2),
                          "plain_ms": round(1e3 * float(iso[idx].sum()), 2)}
    idx = np.flatnonzero(nvs >= 15)
    dt, _ = run_engine_pass(eng, srs, [specs[i] for i in idx], [bufs[i] for i in idx])
    by_class["rows_2e15_2e22"] = {"shards": int(len(idx)), "engine_ms": round(1e3 * dt, 2),
                                  "plain_ms": round(1e3 * float(iso[idx].sum()), 2)}
    value = len(specs) / t_eng
    log("single large shard (2^20 x 8) with kernel timing")
    single, roof_single, _ = run_single(ctx, srs, args)
    roof = roofline(timed, "k_accumulate (MSM bucket accumulation), the engine's timed passes")
    everything = roofline(stats_sum(stats_sum(ctx.stats(), warm), timed), "all launches of the process")
    roof["process"] = {"launches": everything["launches"], "avg_launch_ms": everything["avg_launch_ms"]}
    roof["alone"] = {k: roof_single[k] for k in ("kernel", "achieved", "frac", "traffic", "launches",
                                                 "avg_launch_ms", "entries_per_launch")}
    # the same multiplies over the whole device and the whole step: every
    # accumulation of a pass against the pass's wall time
    entries_pass = timed["accumulate_entries"] / max(1, args.steps)
    dev_rate = IMUL_PER_MADD * entries_pass / t_eng
    roof["device_wide"] = {"imul_per_pass": round(IMUL_PER_MADD * entries_pass),
                           "achieved": round(dev_rate / 1e12, 3), "frac": round(dev_rate / PEAK_IMUL, 4)}
    eng.close()
    log("uniform-scalar MSM at 2^22 points")
    uni = msm_uniform_line(ctx, srs, args.seed)
    log(f"uniform 2^22 MSM: {uni['ms']} ms, {uni['g1_points_per_s'] / 1e6:.1f} M points/s")
    log("served stream: Poisson arrivals at 0.5x, 1x, 1.5x the serving engine's capacity")
    seng = apg.Engine(0, workers=SERVE_WORKERS, hbm_budget=200 << 30, lane_workers=SERVE_RESERVE,
                      lane_cus=SERVE_LANE_CUS, small_scalars=args.small_scalars)
    seng.set_policy("shortest", SERVE_RESERVE, args.small_scalars)
    seng.set_preemption(SERVE_PREEMPT, SERVE_PREEMPT_MODE)
    sargs = argparse.Namespace(**{**vars(args), "workers": SERVE_WORKERS, "policy": "shortest",
                                  "reserve": SERVE_RESERVE, "lane_cus": SERVE_LANE_CUS, "preempt": SERVE_PREEMPT,
                                  "preempt_mode": SERVE_PREEMPT_MODE,
                                  "warmup": 1})
    served = serve(seng, srs, specs, bufs, plain, iso, sargs)
    # the same arrival rates into the urgency order (the alternative for
    # loads near and above capacity)
    seng.set_preemption(SERVE_URGENCY_RATIO, "urgency")
    uargs = argparse.Namespace(**{**vars(sargs), "preempt": SERVE_URGENCY_RATIO, "preempt_mode": "urgency",
                                  "serve_capacity": served["capacity_shards_per_s"], "warmup": 0})
    log("served stream, urgency order, at the same rates")
    served_u = serve(seng, srs, specs, bufs, plain, iso, uargs)
    seng.close()
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "shards/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * t_eng, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32x8 (BN254 Fr/Fq Montgomery, exact)",
        "data": "synthetic (counter-mode ChaCha20 witness mix in HBM; SRS [tau^i]G generated on device)",
        "config": {
            "workload": f"mixed batch: {len(specs)} shards, 23 families of 4-24 columns, log2 rows 60% 8-14, "
                        f"30% 15-19, 10% 20-22; commit MSMs + degree-3 batched sumcheck per shard",
            "shards": len(specs), "scalars": int(sum(s.elements for s in specs)),
            "engine_workers": args.workers, "engine_policy": args.policy, "parallelism": "1 GPU",
        },
        "plain_shards_per_s": round(len(specs) / t_plain, 3),
        "batched_over_plain": round(value / (len(specs) / t_plain), 3),
        "by_class": by_class,
        "bit_exact_vs_plain": bool(exact),
        "hbm": {"srs_gib": round(srs.device_bytes() / 2**30, 3),
                "columns_gib": round(sum(32 * s.elements for s in specs) / 2**30, 3),
                "engine_arena_gib": round(info["arena_bytes"] / 2**30, 3)},
        "single_shard": single,
        "g1_msm_points_per_s": single["g1_msm_points_per_s"],
        "msm_uniform_2e22": uni,
        "serve": {"engine": f"{SERVE_WORKERS} workers, {SERVE_RESERVE} of them kept for shards of at most 2^20 "
                            f"scalars" + (f" on {SERVE_LANE_CUS} CUs of their own" if SERVE_LANE_CUS else "") +
                            f", shortest first; a job pauses at phase boundaries while jobs "
                            f"{SERVE_PREEMPT:g}x smaller run; 200 GB admission budget",
                  "pauses": served["pauses"],
                  "capacity_shards_per_s": served["capacity_shards_per_s"], "rows": served["serve"],
                  "bit_exact_vs_plain": served["bit_exact_vs_plain"],
                  "urgency_order": {"preempt": SERVE_URGENCY_RATIO, "pauses": served_u["pauses"] - served["pauses"],
                                    "rows": served_u["serve"],
                                    "bit_exact_vs_plain": served_u["bit_exact_vs_plain"]}},
        "roofline": roof,
        "cpu_baseline": None,
        "parity": None,
    }
    if args.cpu_budget > 0:
        log("CPU baseline: oracle on a sample of the batch")
        cb, ok, nsample = cpu_baseline(srs, specs, bufs, plain, args.cpu_budget)
        out["cpu_baseline"] = cb
        out["parity"] = bool(ok and exact and served["bit_exact_vs_plain"] and served_u["bit_exact_vs_plain"])
        log(f"CPU baseline {cb['value']} shards/s over {nsample} sampled shards, parity {ok}")
    print(json.dumps(out), flush=True)


def serve(eng, srs, specs, bufs, plain, iso, args):
    """Poisson arrivals at 0.5x, 1x, 1.5x the batched capacity."""
    eng.set_profiling(False)
    for _ in range(max(0, args.warmup)):  # workers' scratch grows to its steady state
        run_engine_pass(eng, srs, specs, bufs)
    dt, _ = run_engine_pass(eng, srs, specs, bufs)  # capacity probe
    capacity = len(specs) / dt
    measured = capacity
    if getattr(args, "serve_capacity", 0.0) > 0:
        capacity = args.serve_capacity
    log(f"batched capacity {capacity:.1f} shards/s")
    iso_ms = iso * 1e3
    rows = []
    exact = True
    for mult in (0.5, 1.0, 1.5):
        rate = mult * capacity
        arrivals = wl.poisson_arrivals(rate, len(specs), seed=args.seed + int(mult * 10))
        order = np.random.default_rng(args.seed).permutation(len(specs))
        t0 = time.perf_counter()
        ids = []
        peak = 0
        for a, i in zip(arrivals, order):
            while time.perf_counter() - t0 < a:
                time.sleep(0.0002)
            s = specs[i]
            ids.append((i, eng.submit_shard(srs, bufs[i], s.num_vars, s.columns, s.degree)))
            peak = max(peak, eng.stats()["hbm_in_use"])
        lats, ratios, detail = [], [], []
        for i, jid in ids:
            p, _, st = eng.wait(jid)
            exact = exact and same(plain[i], p)
            lats.append(st["total_ms"])
            ratios.append(st["total_ms"] / iso_ms[i])
            detail.append((ratios[-1], i, st))
        wall = time.perf_counter() - t0
        lats, ratios = np.array(lats), np.array(ratios)
        nv_of = np.array([specs[i].num_vars for i, _ in ids])
        rows.append({
            "offered_rate": round(rate, 2), "load": mult, "shards_per_s": round(len(specs) / wall, 2),
            "latency_ms": {q: round(float(np.percentile(lats, p)), 3) for q, p in
                           (("p50", 50), ("p99", 99), ("p999", 99.9))},
            "latency_over_isolated": {q: round(float(np.percentile(ratios, p)), 3) for q, p in
                                      (("p50", 50), ("p99", 99))},
            # the same by size class of the mixed batch: 2^8-2^14, 2^15-2^19, 2^20-2^22 rows
            "over_isolated_by_class": {
                name: {"shards": int(sel.sum()),
                       "p50": round(float(np.percentile(ratios[sel], 50)), 3),
                       "p99": round(float(np.percentile(ratios[sel], 99)), 3),
                       "max": round(float(ratios[sel].max()), 3)}
                for name, sel in (("rows_2e8_2e14", nv_of <= 14), ("rows_2e15_2e19", (nv_of >= 15) & (nv_of <= 19)),
                                  ("rows_2e20_2e22", nv_of >= 20)) if sel.any()},
            "hbm_admitted_peak_gib": round(peak / 2**30, 3),
            "worst": [{"rows_log2": specs[i].num_vars, "columns": specs[i].columns,
                       "isolated_ms": round(float(iso_ms[i]), 3), "queue_ms": round(st["queue_ms"], 3),
                       "run_ms": round(st["run_ms"], 3), "worker": st.get("worker"), "over_isolated": round(r, 2)}
                      for r, i, st in sorted(detail, key=lambda t: -t[0])[:5]],
        })
        log(f"load {mult}: {rows[-1]['shards_per_s']} shards/s, p99 {rows[-1]['latency_ms']['p99']} ms")
    info = eng.stats()
    return {"metric": METRIC, "value": rows[1]["shards_per_s"], "unit": "shards/s", "n_gpus": 1,
            "steps": 1, "warmup": 1, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "u32x8 (BN254 Fr/Fq Montgomery, exact)",
            "data": "synthetic, seeded Poisson arrivals",
            "pauses": eng.pauses(),
            "config": {"workload": "served stream of the mixed batch's shards", "parallelism": "1 GPU",
                       "engine_workers": args.workers, "engine_policy": args.policy, "preempt": args.preempt,
                       "preempt_mode": args.preempt_mode,
                       "reserve": args.reserve, "lane_cus": args.lane_cus, "small_scalars": args.small_scalars, "hbm_budget_gib": 200},
            "capacity_shards_per_s": round(capacity, 2), "measured_capacity_shards_per_s": round(measured, 2),
            "serve": rows, "bit_exact_vs_plain": bool(exact),
            "hbm": {"srs_gib": round(srs.device_bytes() / 2**30, 3),
                    "engine_arena_gib": round(info["arena_bytes"] / 2**30, 3)}}


if __name__ == "__main__":
    main()
