"""The Poseidon2 transcript on the device (apg_poseidon2_permute_dev,
apg_transcript_run_dev): the reference's permutation vectors
(crates/transcript/tests/vectors/poseidon2_perm.txt) and its transcript
scripts (transcript_cases.txt) replayed on the device, one state or
transcript per thread, and random programs against the host transcript."""
import importlib

import numpy as np
import pytest

from conftest import golden_lines
from test_oracle_poly_transcript import TAGS

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def apg():
    m = importlib.import_module("apogee-vm_amd")
    m.default_context()
    return m


def test_device_permutation_matches_reference_vectors(apg, orc):
    ins, want = [], []
    for rec in golden_lines("poseidon2_perm.txt"):
        ins.append(np.stack([orc.from_hex_le(orc.FR, h) for h in rec[1:4]]))
        want.append(rec[4:7])
    out = apg.poseidon2_permute_device(np.stack(ins))
    got = [[orc.to_hex_le(orc.FR, out[i, l]) for l in range(3)] for i in range(len(ins))]
    assert len(ins) >= 100 and got == want
    # edge states and random ones against the oracle, in one launch of 4 + 600
    pm1 = orc.neg(orc.FR, orc.from_u64(orc.FR, 1))
    rng = orc.SplitMix(77)
    edge = [orc.fe(), orc.from_u64(orc.FR, 1), pm1, orc.neg(orc.FR, orc.from_u64(orc.FR, 2))]
    states = [np.stack([edge[(i + l) % 4] for l in range(3)]) for i in range(4)]
    states += [np.stack([rng.next_fr() for _ in range(3)]) for _ in range(600)]
    out = apg.poseidon2_permute_device(np.stack(states))
    for i, s in enumerate(states):
        assert np.array_equal(out[i], orc.poseidon2_permute(s)), i


def test_device_transcript_replays_reference_cases(apg, orc):
    """Every case of transcript_cases.txt as one device program, all cases in
    one launch: each sample and challenge equals the reference's, and each
    transcript ends in the state the host transcript reaches."""
    progs, expect, hosts = [], [], []
    for rec in golden_lines("transcript_cases.txt"):
        op = rec[0]
        if op == "case":
            t = apg.Transcript()
            prog, exp, host = apg.TranscriptProgram(t), [], apg.Transcript()
        elif op == "end":
            progs.append(prog)
            expect.append(exp)
            hosts.append(host)
        elif op == "observe":
            x = orc.from_hex_le(orc.FR, rec[1])
            prog.observe(x)
            host.observe(x)
        elif op == "sample":
            prog.sample()
            exp.append(rec[1])
            host.sample()
        elif op in ("append_scalar", "append_scalars"):
            xs = [orc.from_hex_le(orc.FR, rec[2])] if op == "append_scalar" else \
                [orc.from_hex_le(orc.FR, h) for h in rec[3:3 + int(rec[2])]]
            prog.append_scalars(TAGS[rec[1]], xs)
            host.append_scalars(TAGS[rec[1]], xs)
        elif op == "append_bytes":
            b = b"" if rec[3] == "-" else bytes.fromhex(rec[3])
            prog.append_bytes(TAGS[rec[1]], b)
            host.append_bytes(TAGS[rec[1]], b)
        elif op == "challenge":
            prog.challenge(TAGS[rec[1]])
            exp.append(rec[2])
            host.challenge_scalar(TAGS[rec[1]])
        else:
            raise AssertionError(op)
    outs = apg.run_transcripts_device(progs)
    assert len(progs) >= 10
    for prog, exp, host, out in zip(progs, expect, hosts, outs):
        assert [orc.to_hex_le(orc.FR, v) for v in out] == exp
        assert prog.transcript.state() == host.state()


def test_device_transcripts_match_host_on_random_programs(apg, orc):
    """300 transcripts with random interleavings of observes, typed messages,
    samples and challenges (odd and even absorb counts, samples after
    samples), seeded differently, against the host transcript."""
    rng = np.random.default_rng(5)
    srng = orc.SplitMix(9)
    progs, hosts, want = [], [], []
    for i in range(300):
        seed = apg.Transcript()
        seed.append_bytes(2, b"seed %d" % i)
        t, h = seed.clone(), seed.clone()
        prog, w = apg.TranscriptProgram(t), []
        for _ in range(int(rng.integers(1, 12))):
            kind = int(rng.integers(0, 5))
            if kind == 0:
                x = srng.next_fr()
                prog.observe(x)
                h.observe(x)
            elif kind == 1:
                prog.sample()
                w.append(h.sample())
            elif kind == 2:
                xs = [srng.next_fr() for _ in range(int(rng.integers(0, 6)))]
                prog.append_scalars(4, xs)
                h.append_scalars(4, xs)
            elif kind == 3:
                prog.challenge(5)
                w.append(h.challenge_scalar(5))
            else:
                data = bytes(rng.integers(0, 256, size=int(rng.integers(0, 70)), dtype=np.uint8))
                prog.append_bytes(2, data)
                h.append_bytes(2, data)
        progs.append(prog)
        hosts.append(h)
        want.append(w)
    outs = apg.run_transcripts_device(progs)
    for prog, h, w, out in zip(progs, hosts, want, outs):
        assert len(out) == len(w) and all(np.array_equal(a, b) for a, b in zip(out, w))
        assert prog.transcript.state() == h.state()


def test_device_transcript_refusals(apg):
    import ctypes as C

    L = apg.lib()
    ctx = apg.default_context()
    t = apg.Transcript()
    hs = (C.c_void_p * 1)(t.h)
    ops = np.array([[0, 5, 0, 0]], dtype=np.uint32)  # operand 5 of 1
    offs = np.array([0, 1], dtype=np.uint32)
    x = np.zeros((1, 4), dtype=np.uint64)
    out = np.zeros((1, 4), dtype=np.uint64)
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    before = t.state()
    assert L.apg_transcript_run_dev(ctx.h, hs, 1, p(ops), 1, p(offs), p(x), 1, p(out), 1) == -2
    ops[0] = [7, 0, 0, 0]  # unknown op
    assert L.apg_transcript_run_dev(ctx.h, hs, 1, p(ops), 1, p(offs), p(x), 1, p(out), 1) == -2
    ops[0] = [2, 3, 0, 0]  # output 3 of 1
    assert L.apg_transcript_run_dev(ctx.h, hs, 1, p(ops), 1, p(offs), p(x), 1, p(out), 1) == -2
    bad = np.full((1, 4), 0xFFFFFFFFFFFFFFFF, dtype=np.uint64)  # not canonical
    ops[0] = [0, 0, 0, 0]
    assert L.apg_transcript_run_dev(ctx.h, hs, 1, p(ops), 1, p(offs), p(bad), 1, p(out), 1) == -2
    offs2 = np.array([0, 2], dtype=np.uint32)  # offsets past the ops
    assert L.apg_transcript_run_dev(ctx.h, hs, 1, p(ops), 1, p(offs2), p(x), 1, p(out), 1) == -2
    assert t.state() == before  # refused programs leave the transcript alone
    st = np.full((1, 3, 4), 0xFFFFFFFFFFFFFFFF, dtype=np.uint64)
    assert L.apg_poseidon2_permute_dev(ctx.h, p(st), 1) == -2
