"""Oracle bind / eq / evaluate (crates/poly/tests/vectors/poly_kats.txt) and
the Poseidon2 transcript (crates/transcript/tests/vectors/*.txt,
crates/pcs/tests/vectors/g1_absorb_kats.txt)."""
import numpy as np

from conftest import golden_lines

TAGS = {
    "PROTOCOL_SUITE": 1, "PUBLIC_INPUTS": 2, "COMMITMENT": 3, "SUMCHECK_ROUND": 4,
    "SUMCHECK_CHALLENGE": 5, "EVALUATION_CLAIM": 6, "PCS_OPENING": 7, "WITNESS_DIGEST": 8,
    "SUMCHECK_FINAL_EVALS": 9, "MERCURY_INSTANCE": 10, "MERCURY_ALPHA": 11, "MERCURY_GAMMA": 12,
    "MERCURY_Z": 13, "BDFG_BATCH": 14, "BDFG_POINT": 15, "PAIRING_MERGE": 16, "MERCURY_BATCH": 17,
    "ACCUMULATOR_DIGEST": 18, "ACCUMULATOR_MERGE": 19, "PUBLIC_INPUT_STREAM": 20,
    "PUBLIC_OUTPUT_STREAM": 21,
}  # crates/constants/src/lib.rs:579-660


def _poly_kats():
    src, ch, binds, eqs, points, results = [], {}, {}, {}, {}, {}
    for rec in golden_lines("poly_kats.txt"):
        k = rec[0]
        if k == "value":
            src.append(int(rec[2], 16))
        elif k == "challenge":
            ch[int(rec[1])] = rec[2]
        elif k == "bind":
            binds.setdefault(int(rec[1]), {})[int(rec[2])] = rec[3]
        elif k == "eq":
            eqs.setdefault(int(rec[1]), {})[int(rec[2])] = rec[3]
        elif k == "point":
            points.setdefault(int(rec[1]), {})[int(rec[2])] = rec[3]
        elif k == "result":
            results[int(rec[1])] = rec[2]
    return src, ch, binds, eqs, points, results


def test_poly_kats(orc):
    src, ch, binds, eqs, points, results = _poly_kats()
    table = np.stack([orc.from_u64(orc.FR, v) for v in src])
    chal = [orc.from_hex_le(orc.FR, ch[i]) for i in range(len(ch))]
    t = table
    for k in range(1, len(chal) + 1):
        t = orc.bind(t, chal[k - 1])
        exp = binds[k]
        assert len(exp) == t.shape[0]
        for i, h in exp.items():
            assert orc.to_hex_le(orc.FR, t[i]) == h, (k, i)
    for k, exp in eqs.items():
        r = np.stack(chal[:k]) if k else orc.fe(0)
        e = orc.eq_table(r)
        assert e.shape[0] == len(exp)
        for i, h in exp.items():
            assert orc.to_hex_le(orc.FR, e[i]) == h, (k, i)
    for c, pt in points.items():
        point = np.stack([orc.from_hex_le(orc.FR, pt[j]) for j in range(len(pt))])
        assert orc.to_hex_le(orc.FR, orc.evaluate(table, point)) == results[c]


def test_eq_table_matches_eq_eval(orc):
    rng = orc.SplitMix(9)
    r = np.stack([rng.next_fr() for _ in range(5)])
    e = orc.eq_table(r)
    one, zero = orc.from_u64(orc.FR, 1), orc.fe()
    for y in range(32):
        pt = np.stack([one if (y >> j) & 1 else zero for j in range(5)])
        assert np.array_equal(e[y], orc.eq_eval(r, pt))


def test_poseidon2_kats(orc):
    n = 0
    for rec in golden_lines("poseidon2_perm.txt"):
        st = np.stack([orc.from_hex_le(orc.FR, h) for h in rec[1:4]])
        out = orc.poseidon2_permute(st)
        assert [orc.to_hex_le(orc.FR, out[i]) for i in range(3)] == rec[4:7]
        n += 1
    assert n >= 100


def replay_cases(orc, transcript_factory):
    """Replays transcript_cases.txt through a transcript built by the factory;
    returns the number of cases."""
    cases = 0
    t = None
    for rec in golden_lines("transcript_cases.txt"):
        op = rec[0]
        if op == "case":
            t = transcript_factory()
        elif op == "end":
            cases += 1
        elif op == "observe":
            t.observe(orc.from_hex_le(orc.FR, rec[1]))
        elif op == "sample":
            assert orc.to_hex_le(orc.FR, t.sample()) == rec[1]
        elif op == "append_scalar":
            t.append_scalars(TAGS[rec[1]], orc.from_hex_le(orc.FR, rec[2]))
        elif op == "append_scalars":
            n = int(rec[2])
            xs = np.stack([orc.from_hex_le(orc.FR, h) for h in rec[3:3 + n]]) if n else orc.fe(0)
            t.append_scalars(TAGS[rec[1]], xs)
        elif op == "append_bytes":
            b = b"" if rec[3] == "-" else bytes.fromhex(rec[3])
            assert len(b) == int(rec[2])
            t.append_bytes(TAGS[rec[1]], b)
        elif op == "challenge":
            assert orc.to_hex_le(orc.FR, t.challenge(TAGS[rec[1]])) == rec[2]
        else:
            raise AssertionError(op)
    return cases


def test_transcript_cases(orc):
    assert replay_cases(orc, orc.Transcript) >= 10


def test_g1_absorb_kats(orc):
    for rec in golden_lines("g1_absorb_kats.txt"):
        if rec[0] == "single":
            limbs = orc.g1_limbs(bytes.fromhex(rec[1]))
            assert [orc.to_hex_le(orc.FR, limbs[i]) for i in range(4)] == rec[2:6]
        elif rec[0] == "pair":
            a = orc.g1_limbs(bytes.fromhex(rec[1]))
            b = orc.g1_limbs(bytes.fromhex(rec[2]))
            got = [orc.to_hex_le(orc.FR, x) for x in list(a) + list(b)]
            assert got == rec[3:11]


def test_chacha20_rfc8439_block():
    # RFC 8439 section 2.3.2 test vector
    import oracle

    key = bytes(range(32))
    nonce = bytes.fromhex("000000090000004a00000000")
    out = oracle.chacha_block(key, 1, nonce)
    assert out.hex() == (
        "10f1e7e4d13b5915500fdd1fa32071c4c7d1f4c733c068030422aa9ac3d46c4e"
        "d2826446079faa0914c2d705d98b02a2b5129cd1de164eb9cbd083e8a2503c4e"
    )
