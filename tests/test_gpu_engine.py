"""The shard engine (continuous batching) through the C ABI: each shard's
bytes are the plain one-shard path's, alone and in the mix, and equal the
oracle's; admission refuses a job above the HBM budget."""
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def apg():
    m = importlib.import_module("apogee-vm_amd")
    m.default_context()
    return m


@pytest.fixture(scope="module")
def setup(apg):
    wl = importlib.import_module("apogee-vm_amd.workload")
    ctx = apg.default_context()
    specs = wl.mixed_batch(seed=7, shards=14, families=5, min_cols=1, max_cols=6,
                           classes=((0.5, 3, 7), (0.5, 8, 11)))
    tau = apg.synthetic(99, 1, "uniform")[0]
    nmax = max(s.rows for s in specs)
    srs = apg.Srs.generate(tau, nmax, ctx)
    bufs = []
    for s in specs:
        b = apg.DeviceBuffer(ctx, 32 * s.elements)
        for c in range(s.columns):
            apg.synthetic_device(ctx, s.seed + c, s.rows, "witness", out=b, offset_elems=c * s.rows)
        bufs.append(b)
    return specs, srs, bufs


def _plain(apg, srs, spec, buf):
    t = apg.Transcript()
    p = apg.prove_shard(srs, buf, spec.num_vars, spec.columns, spec.degree, t)
    return p, t.state()


def test_engine_matches_plain_path_in_the_mix(apg, setup):
    specs, srs, bufs = setup
    plain = [_plain(apg, srs, s, b) for s, b in zip(specs, bufs)]
    eng = apg.Engine(0, workers=4, hbm_budget=8 << 30)
    ids = [eng.submit_shard(srs, b, s.num_vars, s.columns, s.degree) for s, b in zip(specs, bufs)]
    for (p, st), jid in zip(plain, ids):
        q, t, stats = eng.wait(jid)
        assert q.commitments == p.commitments
        assert np.array_equal(q.rounds, p.rounds) and np.array_equal(q.final_evals, p.final_evals)
        assert t.state() == st
        assert stats["total_ms"] >= stats["run_ms"] >= 0
    # alone: one shard through a one-worker engine gives the same bytes again
    solo = apg.Engine(0, workers=1, hbm_budget=8 << 30)
    q, t, _ = solo.wait(solo.submit_shard(srs, bufs[3], specs[3].num_vars, specs[3].columns, specs[3].degree))
    assert q.commitments == plain[3][0].commitments and t.state() == plain[3][1]
    info = eng.stats()
    # idle workers are charged the scratch they keep, nothing more
    assert info["completed"] == len(specs) and info["running"] == 0
    assert info["hbm_in_use"] == info["arena_bytes"] > 0 and info["hbm_peak"] >= info["hbm_in_use"]


def test_engine_lanes_match_plain_path(apg, setup):
    """A fast lane on its own CUs (short jobs only, shortest first) beside
    the other workers: every shard's bytes are the plain path's, whichever
    lane and queue order ran it."""
    specs, srs, bufs = setup
    plain = [_plain(apg, srs, s, b) for s, b in zip(specs, bufs)]
    small = float(np.median([s.elements for s in specs]))
    eng = apg.Engine(0, workers=3, hbm_budget=8 << 30, lane_workers=1, lane_cus=32, small_scalars=small)
    ids = [eng.submit_shard(srs, b, s.num_vars, s.columns, s.degree) for s, b in zip(specs, bufs)]
    workers = set()
    for (p, st), jid in zip(plain, ids):
        q, t, stats = eng.wait(jid)
        assert q.commitments == p.commitments and np.array_equal(q.rounds, p.rounds)
        assert t.state() == st
        workers.add(stats["worker"])
    assert 0 in workers  # the fast lane ran some of them
    eng.set_policy("fifo", 1, small)
    q, _, _ = eng.wait(eng.submit_shard(srs, bufs[0], specs[0].num_vars, specs[0].columns, specs[0].degree))
    assert q.commitments == plain[0][0].commitments
    # longest first, no lane: the same bytes in any completion order
    eng.set_policy("longest", 0, 0.0)
    ids = [eng.submit_shard(srs, b, s.num_vars, s.columns, s.degree) for s, b in zip(specs, bufs)]
    for (p, st), jid in zip(plain, ids):
        q, t, _ = eng.wait(jid)
        assert q.commitments == p.commitments and np.array_equal(q.rounds, p.rounds) and t.state() == st
    with pytest.raises(ValueError):
        eng.set_policy("widest")


def test_engine_matches_oracle(apg, orc, setup):
    specs, srs, bufs = setup
    s, b = specs[-1], bufs[-1]
    eng = apg.Engine(0, workers=2, hbm_budget=8 << 30)
    q, _, _ = eng.wait(eng.submit_shard(srs, b, s.num_vars, s.columns, s.degree))
    cols = b.download().reshape(s.columns, s.rows, 4)
    cms, rounds, fin = orc.shard_script(srs.g1(0, s.rows), cols, s.degree, orc.Transcript())
    assert q.commitments == cms and np.array_equal(q.rounds, rounds) and np.array_equal(q.final_evals, fin)


def test_engine_seeded_transcript_and_rejection(apg, setup):
    specs, srs, bufs = setup
    s, b = specs[0], bufs[0]
    seed = apg.Transcript()
    seed.append_bytes(2, b"shard seed")
    t = seed.clone()
    p = apg.prove_shard(srs, b, s.num_vars, s.columns, s.degree, t)
    eng = apg.Engine(0, workers=2, hbm_budget=8 << 30)
    q, t2, _ = eng.wait(eng.submit_shard(srs, b, s.num_vars, s.columns, s.degree, seed=seed))
    assert q.commitments == p.commitments and t2.state() == t.state()
    tiny = apg.Engine(0, workers=1, hbm_budget=1 << 20)
    jid = tiny.submit_shard(srs, b, s.num_vars, s.columns, s.degree)
    with pytest.raises(apg.ApgError):
        tiny.wait(jid)
    # the refusal leaves the engine serving
    ok = apg.Engine(0, workers=1, hbm_budget=8 << 30)
    q2, _, _ = ok.wait(ok.submit_shard(srs, b, s.num_vars, s.columns, s.degree))
    assert q2.commitments == p.commitments


def test_engine_tables_outgrown_while_in_use(apg, orc):
    """Shards of one window width (2^17..2^18 rows share c = 17) submitted
    smallest first to a fresh SRS with no prepared tables: a longer shard
    replaces the width's table while shorter shards' MSMs still read the old
    one, which must stay alive until they finish. Every result equals the
    plain path's, run afterwards on the grown table."""
    ctx = apg.default_context()
    assert apg.window_bits(1 << 17) == apg.window_bits(1 << 18)
    tau = apg.synthetic(4321, 1, "uniform")[0]
    srs = apg.Srs.generate(tau, 1 << 18, ctx)
    specs = [(17, 2, 900 + i) if i < 6 else (18, 2, 900 + i) for i in range(9)]
    bufs = []
    for nv, k, seed in specs:
        b = apg.DeviceBuffer(ctx, 32 * k << nv)
        for c in range(k):
            apg.synthetic_device(ctx, seed * 10 + c, 1 << nv, "witness", out=b, offset_elems=c << nv)
        bufs.append(b)
    eng = apg.Engine(0, workers=4, hbm_budget=16 << 30)
    ids = [eng.submit_shard(srs, b, nv, k, 3) for (nv, k, _), b in zip(specs, bufs)]
    got = [eng.wait(j)[0] for j in ids]
    eng.close()
    for (nv, k, _), b, q in zip(specs, bufs, got):
        p = apg.prove_shard(srs, b, nv, k, 3, apg.Transcript())
        assert q.commitments == p.commitments and np.array_equal(q.rounds, p.rounds)
    # and one of each size against the oracle MSM
    for i in (0, len(specs) - 1):
        nv, k, _ = specs[i]
        cols = bufs[i].download().reshape(k, 1 << nv, 4)
        assert got[i].commitments[0] == orc.jac_to_wire(orc.msm(srs.g1(0, 1 << nv), cols[0], threads=8, workers=8))
    srs.release()


@pytest.mark.parametrize("mode", ["size", "urgency"])
def test_engine_preemption_matches_plain_path(apg, mode):
    """Phase-boundary preemption: a long shard (2^18 rows x 8) whose
    accumulation runs in chunks with yield points between them and between
    its sumcheck rounds, and short shards submitted while it runs, which make
    it pause (by size: at least once; by urgency: as the urgencies fall).
    Every shard's bytes equal the plain path's, the long one's included."""
    import time

    ctx = apg.default_context()
    tau = apg.synthetic(2468, 1, "uniform")[0]
    srs = apg.Srs.generate(tau, 1 << 18, ctx)
    specs = [(18, 8, 70)] + [(10 + (i % 3), 3 + (i % 4), 71 + i) for i in range(12)]
    bufs = []
    for nv, k, seed in specs:
        b = apg.DeviceBuffer(ctx, 32 * k << nv)
        for c in range(k):
            apg.synthetic_device(ctx, seed * 10 + c, 1 << nv, "witness", out=b, offset_elems=c << nv)
        bufs.append(b)
    plain = []
    for (nv, k, _), b in zip(specs, bufs):
        t = apg.Transcript()
        plain.append((apg.prove_shard(srs, b, nv, k, 3, t), t.state()))
    eng = apg.Engine(0, workers=4, hbm_budget=16 << 30)
    eng.set_policy("shortest", 0, 0.0)
    eng.set_preemption(1.0, mode)
    ids = [eng.submit_shard(srs, bufs[0], specs[0][0], specs[0][1], 3)]
    t0 = time.perf_counter()
    while eng.stats()["running"] == 0 and time.perf_counter() - t0 < 5:
        time.sleep(0.0002)
    ids += [eng.submit_shard(srs, b, nv, k, 3) for (nv, k, _), b in zip(specs[1:], bufs[1:])]
    for (p, st), jid in zip(plain, ids):
        q, t, _ = eng.wait(jid)
        assert q.commitments == p.commitments
        assert np.array_equal(q.rounds, p.rounds) and np.array_equal(q.final_evals, p.final_evals)
        assert t.state() == st
    if mode == "size":
        assert eng.pauses() >= 1
    # preemption off again: nothing more pauses
    before = eng.set_preemption(0.0)
    ids = [eng.submit_shard(srs, b, nv, k, 3) for (nv, k, _), b in zip(specs, bufs)]
    for (p, _), jid in zip(plain, ids):
        assert eng.wait(jid)[0].commitments == p.commitments
    assert eng.pauses() == before
    eng.close()
    srs.release()
