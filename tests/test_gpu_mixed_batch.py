"""The benchmark's headline workloads on the GPU, at full size.

Config 4, the mixed batch: 207 shards from 23 families of 4-24 witness-like
columns, 2^8-2^22 rows (apogee-vm_amd/workload.py), proved by the shard
engine (12 workers) and by the plain one-shard-per-call path:
  * every shard's commitments, rounds, final evaluations and final
    transcript state are the same on both paths;
  * every shard of at most 2^18 rows equals the oracle's whole shard script
    on the downloaded columns;
  * every larger shard (2^19-2^22 rows) passes the sumcheck verifier's
    identities against values the oracle computes from its own copy of the
    inputs (regenerated from the seeds): the transcript replayed from the
    proof's messages reaches the proof's final state; g_0(0) + g_0(1) equals
    the claim sum_x eq(r, x) t_0(x) t_1(x); each round's g_j(0) + g_j(1)
    equals g_{j-1}(rho_{j-1}); the last round at its challenge equals
    eq(r, rho) t_0(rho) t_1(rho); and the final evaluations equal the
    oracle's combined tables evaluated at rho;
  * commitments of the large shards against the oracle MSM: every column of
    one 2^20 shard, and a random linear combination of every column of one
    2^21 and one 2^22 shard (sum_c g^c cm_c = MSM(bases, sum_c g^c col_c)).
Config 5, the served stream: the same shards arriving as a seeded Poisson
process at 0.5x, 1x and 1.5x the batch's measured capacity into the serving
engine of the benchmark (6 workers, 2 of them kept for short shards, shortest
first, jobs pausing at phase boundaries for jobs 2x smaller): every result
equals the plain path's, so a shard's bytes are the same alone, in the mix and
when paused and resumed.
The reference tests determinism under concurrency the same way
(crates/prover/tests/streaming.rs:54-110).
"""
import concurrent.futures as cf
import importlib
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEED = 20261203  # the benchmark's default seed
SMALL_MAX = 18  # log2 rows of the shards checked against the whole oracle script
THREADS = 16


@pytest.fixture(scope="module")
def apg():
    m = importlib.import_module("apogee-vm_amd")
    m.default_context()
    return m


@pytest.fixture(scope="module")
def batch(apg):
    wl = importlib.import_module("apogee-vm_amd.workload")
    ctx = apg.default_context()
    specs = wl.mixed_batch(seed=SEED)
    assert len(specs) == 207 and max(s.num_vars for s in specs) == 22 and min(s.num_vars for s in specs) == 8
    tau = apg.synthetic(SEED, 1, "uniform")[0]
    srs = apg.Srs.generate(tau, 1 << 22, ctx)
    for nv in sorted({s.num_vars for s in specs}, reverse=True):
        srs.prepare(1 << nv)
    bufs = []
    for s in specs:
        b = apg.DeviceBuffer(ctx, 32 * s.elements)
        for c in range(s.columns):
            apg.synthetic_device(ctx, s.seed + c, s.rows, "witness", out=b, offset_elems=c * s.rows)
        bufs.append(b)
    ctx.synchronize()
    plain = []
    for s, b in zip(specs, bufs):
        t = apg.Transcript()
        p = apg.prove_shard(srs, b, s.num_vars, s.columns, s.degree, t)
        plain.append((p, t.state()))
    eng = apg.Engine(0, workers=12, hbm_budget=200 << 30)
    t0 = time.perf_counter()
    ids = [eng.submit_shard(srs, b, s.num_vars, s.columns, s.degree) for s, b in zip(specs, bufs)]
    engine = [eng.wait(j) for j in ids]
    capacity = len(specs) / (time.perf_counter() - t0)
    eng.close()
    print(f"mixed batch: 207 shards through 12 engine workers at {capacity:.1f} shards/s", flush=True)
    yield {"specs": specs, "srs": srs, "bufs": bufs, "plain": plain, "engine": engine, "capacity": capacity}
    for b in bufs:
        b.free()
    srs.release()


def _same(p, q):
    return p.commitments == q.commitments and np.array_equal(p.rounds, q.rounds) and np.array_equal(
        p.final_evals, q.final_evals)


def test_engine_equals_plain_on_every_shard(batch):
    bad = [s.index for s, (p, st), (q, t, _) in zip(batch["specs"], batch["plain"], batch["engine"])
           if not (_same(p, q) and t.state() == st)]
    assert not bad, f"shards differing between the engine and the plain path: {bad}"


def test_device_inputs_are_the_oracle_inputs(apg, orc, batch):
    """The large-shard checks regenerate columns on the host from the seeds;
    the device columns are those same values (first and last column of the
    smallest and the largest shard)."""
    specs, bufs = batch["specs"], batch["bufs"]
    for i in (min(range(len(specs)), key=lambda i: specs[i].rows), max(range(len(specs)), key=lambda i: specs[i].rows)):
        s = specs[i]
        cols = bufs[i].download().reshape(s.columns, s.rows, 4)
        for c in (0, s.columns - 1):
            assert np.array_equal(cols[c], orc.fill_ctr(s.seed + c, s.rows, "witness")), (i, c)


def test_small_shards_match_oracle_script(orc, batch):
    specs, bufs, srs = batch["specs"], batch["bufs"], batch["srs"]
    small = [i for i, s in enumerate(specs) if s.num_vars <= SMALL_MAX]
    assert len(small) == 161
    bases = srs.g1(0, 1 << SMALL_MAX)
    cols = {i: bufs[i].download().reshape(specs[i].columns, specs[i].rows, 4) for i in small}

    def one(i):
        s = specs[i]
        cms, rounds, fin = orc.shard_script(bases[:s.rows], cols[i], s.degree, orc.Transcript(), workers=2)
        p = batch["plain"][i][0]
        return i, p.commitments == cms and np.array_equal(p.rounds, rounds) and np.array_equal(p.final_evals, fin)

    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(THREADS // 2) as ex:
        results = list(ex.map(one, sorted(small, key=lambda i: -specs[i].elements)))
    bad = sorted(i for i, ok in results if not ok)
    print(f"oracle shard script on {len(small)} shards in {time.perf_counter() - t0:.1f} s", flush=True)
    assert not bad, f"shards differing from the oracle: {bad}"


def _horner(orc, coeffs, x):
    acc = coeffs[-1].copy()
    for c in coeffs[-2::-1]:
        acc = orc.add(orc.FR, orc.mul(orc.FR, acc, x), c)
    return acc


def _large_shard_identities(orc, spec, proof, state):
    """The sumcheck verifier's identities for one shard, with the claim and
    the final evaluations computed by the oracle from regenerated inputs.
    Returns the names of the identities that fail."""
    nv, k, f = spec.num_vars, spec.columns, spec.degree - 1
    n = 1 << nv
    failed = []
    t = orc.Transcript()
    t.append_g1(orc.TAG_COMMITMENT, proof.commitments)
    rho_w = t.challenge(orc.TAG_MERCURY_BATCH)
    r = np.stack([t.challenge(orc.TAG_SUMCHECK_CHALLENGE) for _ in range(nv)])
    chal = []
    for j in range(nv):
        t.append_scalars(orc.TAG_SUMCHECK_ROUND, proof.rounds[j])
        chal.append(t.challenge(orc.TAG_SUMCHECK_CHALLENGE))
    chal = np.stack(chal)
    t.append_scalars(orc.TAG_SUMCHECK_FINAL_EVALS, proof.final_evals)
    if t.snapshot()[:96] != state[:96]:
        failed.append("transcript replay")
    # the combined tables t_o = sum_c rho^(o k + c) col_c, one column at a time
    pw = [orc.from_u64(orc.FR, 1)]
    for _ in range(f * k - 1):
        pw.append(orc.mul(orc.FR, pw[-1], rho_w))
    tabs = [orc.fe(n) for _ in range(f)]
    for c in range(k):
        col = orc.fill_ctr(spec.seed + c, n, "witness")
        for o in range(f):
            orc.axpy(pw[o * k + c], col, tabs[o])
        del col
    zero, one = orc.fe(), orc.from_u64(orc.FR, 1)
    g = proof.rounds
    if not np.array_equal(orc.add(orc.FR, _horner(orc, g[0], zero), _horner(orc, g[0], one)),
                          orc.product_claim(r, tabs)):
        failed.append("claim")
    for j in range(1, nv):
        lhs = orc.add(orc.FR, _horner(orc, g[j], zero), _horner(orc, g[j], one))
        if not np.array_equal(lhs, _horner(orc, g[j - 1], chal[j - 1])):
            failed.append(f"round {j}")
    last = orc.eq_eval(r, chal)
    for o in range(f):
        last = orc.mul(orc.FR, last, proof.final_evals[o])
    if not np.array_equal(_horner(orc, g[nv - 1], chal[nv - 1]), last):
        failed.append("final identity")
    for o in range(f):
        if not np.array_equal(orc.evaluate(tabs[o], chal), proof.final_evals[o]):
            failed.append(f"final evaluation {o}")
    return failed


def test_large_shards_verifier_identities(orc, batch):
    specs = batch["specs"]
    large = [i for i, s in enumerate(specs) if s.num_vars > SMALL_MAX]
    assert len(large) == 46 and sum(specs[i].num_vars >= 20 for i in large) == 31

    def one(i):
        p, st = batch["plain"][i]
        return i, _large_shard_identities(orc, specs[i], p, st)

    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(THREADS) as ex:
        results = list(ex.map(one, sorted(large, key=lambda i: -specs[i].elements)))
    print(f"verifier identities on {len(large)} shards in {time.perf_counter() - t0:.1f} s", flush=True)
    bad = {i: f for i, f in results if f}
    assert not bad, f"failed identities: {bad}"


def _wire_to_jac(orc, w):
    a = orc.affine_from_wire(w)
    assert a is not None
    return orc.jac_from_affine(a)


def test_large_shard_commitments_match_oracle_msm(apg, orc, batch):
    specs, bufs, srs = batch["specs"], batch["bufs"], batch["srs"]
    by = lambda nv: min((i for i, s in enumerate(specs) if s.num_vars == nv), key=lambda i: specs[i].columns)
    # every column of the narrowest 2^20 shard
    i20 = by(20)
    s = specs[i20]
    bases = srs.g1(0, s.rows)
    cols = bufs[i20].download().reshape(s.columns, s.rows, 4)
    p = batch["plain"][i20][0]
    for c in range(s.columns):
        assert p.commitments[c] == orc.jac_to_wire(orc.msm(bases, cols[c], threads=8, workers=THREADS)), c
    # a random linear combination of every column of the narrowest 2^21 and 2^22 shards
    for nv in (21, 22):
        i = by(nv)
        s = specs[i]
        bases = srs.g1(0, s.rows)
        gam = orc.fill_uniform(SEED + nv, 1)[0]
        comb = orc.fe(s.rows)
        lhs = None
        w = orc.from_u64(orc.FR, 1)
        p = batch["plain"][i][0]
        for c in range(s.columns):
            orc.axpy(w, orc.fill_ctr(s.seed + c, s.rows, "witness"), comb)
            term = orc.jac_mul(_wire_to_jac(orc, p.commitments[c]), w)
            lhs = term if lhs is None else orc.jac_add(lhs, term)
            w = orc.mul(orc.FR, w, gam)
        rhs = orc.msm(bases, comb, threads=8, workers=THREADS)
        assert orc.jac_to_wire(lhs) == orc.jac_to_wire(rhs), nv


def test_served_stream_matches_plain(apg, batch):
    """Config 5: Poisson arrivals at 0.5x, 1x and 1.5x the batch capacity into
    the serving engine; every result is the plain path's."""
    wl = importlib.import_module("apogee-vm_amd.workload")
    specs, bufs, srs, plain = batch["specs"], batch["bufs"], batch["srs"], batch["plain"]
    eng = apg.Engine(0, workers=6, hbm_budget=200 << 30, lane_workers=2, lane_cus=0, small_scalars=float(1 << 20))
    eng.set_policy("shortest", 2, float(1 << 20))
    eng.set_preemption(2.0)
    order = np.random.default_rng(SEED).permutation(len(specs))
    for load in (0.5, 1.0, 1.5):
        arrivals = wl.poisson_arrivals(load * batch["capacity"], len(specs), seed=SEED + int(load * 10))
        t0 = time.perf_counter()
        ids = []
        for a, i in zip(arrivals, order):
            while time.perf_counter() - t0 < a:
                time.sleep(0.0002)
            s = specs[i]
            ids.append((i, eng.submit_shard(srs, bufs[i], s.num_vars, s.columns, s.degree)))
        lat = []
        for i, jid in ids:
            q, t, st = eng.wait(jid)
            p, state = plain[i]
            assert _same(p, q) and t.state() == state, (load, int(i))
            lat.append(st["total_ms"])
        print(f"served at {load}x: p50 {np.percentile(lat, 50):.2f} ms, p99 {np.percentile(lat, 99):.2f} ms",
              flush=True)
    info = eng.stats()
    assert info["completed"] == 3 * len(specs) and info["queued"] == 0 and info["running"] == 0
    print(f"pauses at phase boundaries: {eng.pauses()}", flush=True)
    assert eng.pauses() > 0  # the long shards gave way to short ones
    eng.close()
