"""Oracle sumcheck against the definition (the reference's own pin:
crates/sumcheck/tests/oracle.rs:45-116, crates/gkr/tests/oracle.rs) and the
round-sum invariants (crates/sumcheck/src/lib.rs:441-459); the commit path
against the committed Mercury commitment (crates/pcs/tests/vectors/mercury_proof.txt)."""
import numpy as np
import pytest

from conftest import golden_lines


def _rand_tables(orc, seed, k, n):
    rng = orc.SplitMix(seed)
    return [np.stack([rng.next_fr() for _ in range(1 << n)]) for _ in range(k)]


def _horner(orc, coeffs, x):
    acc = orc.fe()
    for c in reversed(list(coeffs)):
        acc = orc.add(orc.FR, orc.mul(orc.FR, acc, x), c)
    return acc


def _gate_value(orc, terms, vals):
    acc = orc.fe()
    for coef, fs in terms:
        v = coef
        for f in fs:
            v = orc.mul(orc.FR, v, vals[f])
        acc = orc.add(orc.FR, acc, v)
    return acc


def _naive_round_value(orc, tables, terms, eq_point, bound, x):
    n = eq_point.shape[0]
    rest = n - len(bound) - 1
    one, zero = orc.from_u64(orc.FR, 1), orc.fe()
    acc = orc.fe()
    for y in range(1 << rest):
        pt = list(bound) + [x] + [one if (y >> j) & 1 else zero for j in range(rest)]
        pt = np.stack(pt)
        vals = [orc.evaluate(t, pt) for t in tables]
        acc = orc.add(orc.FR, acc, orc.mul(orc.FR, orc.eq_eval(eq_point, pt), _gate_value(orc, terms, vals)))
    return acc


def _gates(orc):
    one = orc.from_u64(orc.FR, 1)
    minus_one = orc.sub(orc.FR, orc.fe(), one)
    three = orc.from_u64(orc.FR, 3)
    return {
        "product": (2, [(one, [0, 1])]),
        "square_minus": (2, [(one, [0, 0]), (minus_one, [1])]),
        "wide": (5, [(one, [0, 1]), (one, [0, 2]), (minus_one, [3, 4])]),
        "linear": (1, [(three, [0])]),
        "triple": (3, [(one, [0, 1, 2])]),
    }


@pytest.mark.parametrize("gate", ["product", "square_minus", "wide", "linear", "triple"])
@pytest.mark.parametrize("n", [1, 2, 3, 4])
def test_prove_sumcheck_matches_definition(orc, gate, n):
    k, terms = _gates(orc)[gate]
    tables = _rand_tables(orc, 100 * n + len(gate), k, n)
    rng = orc.SplitMix(7 + n)
    eq_point = np.stack([rng.next_fr() for _ in range(n)])
    t = orc.Transcript()
    rounds, point, finals = orc.prove_sumcheck(eq_point, terms, tables, t)
    d = rounds.shape[1] - 1
    assert d == (4 if gate == "triple" else 3)
    claim = None
    for i in range(n):
        for node in range(d + 1):
            x = orc.from_u64(orc.FR, node)
            want = _naive_round_value(orc, tables, terms, eq_point, point[:i], x)
            assert np.array_equal(_horner(orc, rounds[i], x), want), (i, node)
        s01 = orc.add(orc.FR, _horner(orc, rounds[i], orc.fe()), _horner(orc, rounds[i], orc.from_u64(orc.FR, 1)))
        if claim is not None:
            assert np.array_equal(s01, claim)
        claim = _horner(orc, rounds[i], point[i])
    # final check: claim = eq(eq_point, point) * G(final evals)
    fe = [f[0] for f in finals]
    assert np.array_equal(claim, orc.mul(orc.FR, orc.eq_eval(eq_point, point), _gate_value(orc, terms, fe)))


def test_zerocheck_on_a_satisfying_witness(orc):
    """square gate x0*x0 - x1 on a witness with x1 = x0^2: rounds sum to zero and
    the verifier's equation holds (crates/sumcheck/src/lib.rs:418-465)."""
    one = orc.from_u64(orc.FR, 1)
    minus_one = orc.sub(orc.FR, orc.fe(), one)
    terms = [(one, [0, 0]), (minus_one, [1])]
    n = 4
    a = _rand_tables(orc, 31, 1, n)[0]
    b = np.stack([orc.mul(orc.FR, v, v) for v in a])
    t = orc.Transcript()
    rounds, finals = orc.prove_zerocheck(terms, [a, b], t)
    # verifier replay
    v = orc.Transcript()
    r = np.stack([v.challenge(orc.TAG_SUMCHECK_CHALLENGE) for _ in range(n)])
    claim = orc.fe()
    pt = []
    for i in range(n):
        s01 = orc.add(orc.FR, _horner(orc, rounds[i], orc.fe()), _horner(orc, rounds[i], one))
        assert np.array_equal(s01, claim)
        v.append_scalars(orc.TAG_SUMCHECK_ROUND, rounds[i])
        c = v.challenge(orc.TAG_SUMCHECK_CHALLENGE)
        claim = _horner(orc, rounds[i], c)
        pt.append(c)
    g = _gate_value(orc, terms, [finals[0], finals[1]])
    assert np.array_equal(claim, orc.mul(orc.FR, orc.eq_eval(r, np.stack(pt)), g))
    v.append_scalars(orc.TAG_SUMCHECK_FINAL_EVALS, finals)
    assert v.snapshot() == t.snapshot()


@pytest.mark.parametrize("d", [1, 2, 3, 4, 5])
def test_interpolation_hits_every_node(orc, d):
    rng = orc.SplitMix(d)
    vals = np.stack([rng.next_fr() for _ in range(d + 1)])
    c = orc.interpolate(vals)
    for node in range(d + 1):
        assert np.array_equal(_horner(orc, c, orc.from_u64(orc.FR, node)), vals[node])


def test_interpolate_cubic_closed_form_equals_newton(orc):
    # the d == 3 closed form (sumcheck/src/lib.rs:294-305) against the general
    # Newton path, reached through a degree-3 polynomial padded to d = 4
    rng = orc.SplitMix(44)
    vals = np.stack([rng.next_fr() for _ in range(4)])
    c3 = orc.interpolate(vals)
    v4 = np.concatenate([vals, _horner(orc, c3, orc.from_u64(orc.FR, 4))[None]])
    c4 = orc.interpolate(v4)
    assert np.array_equal(c4[:4], c3) and not c4[4].any()


def test_witness_digest_is_a_sponge_over_cells(orc):
    tables = _rand_tables(orc, 5, 2, 3)
    d1 = orc.witness_digest(tables)
    tables[1][5, 0] ^= 1
    assert not np.array_equal(d1, orc.witness_digest(tables))


@pytest.mark.parametrize("k,n", [(1, 0), (2, 3), (5, 4)])
def test_witness_digest_composes_pinned_transcript_calls(orc, k, n):
    """witness_digest (crates/sumcheck/src/lib.rs:228-256) is a fresh sponge
    that absorbs [k, n] as one WITNESS_DIGEST message, then one message per
    column, then squeezes raw. The transcript's absorb/squeeze and the tag
    value are pinned by the transcript known answers, so this pins the digest
    up to the order and framing it states; a dropped header, a swapped count
    or one message for all cells fails here."""
    tables = _rand_tables(orc, 40 + k, k, n)
    t = orc.Transcript()
    t.append_scalars(orc.TAG_WITNESS_DIGEST,
                     np.stack([orc.from_u64(orc.FR, k), orc.from_u64(orc.FR, n)]))
    for c in tables:
        t.append_scalars(orc.TAG_WITNESS_DIGEST, c)
    assert np.array_equal(orc.witness_digest(tables), t.sample())


def test_witness_digest_header_pins_the_shape(orc):
    """The same cells read as one column of 2^(n+1) or two of 2^n digest
    differently: the header carries both counts (crates/sumcheck/src/lib.rs:233-236)."""
    two = _rand_tables(orc, 9, 2, 3)
    one = [np.concatenate(two)]
    assert not np.array_equal(orc.witness_digest(two), orc.witness_digest(one))
    # and swapping the columns moves the digest
    assert not np.array_equal(orc.witness_digest(two), orc.witness_digest(two[::-1]))


def test_mercury_commitment_kat(orc):
    """pcs::commit over the toy SRS (tau = 0xabcdef) equals the committed
    commitment line of mercury_proof.txt (crates/pcs/src/lib.rs:237-247)."""
    recs = list(golden_lines("mercury_proof.txt"))
    tau = orc.from_hex_le(orc.FR, next(r for r in recs if r[0] == "tau")[1])
    nv = int(next(r for r in recs if r[0] == "numvars")[1])
    vals = {int(r[1]): r[2] for r in recs if r[0] == "value"}
    f = np.stack([orc.from_hex_le(orc.FR, vals[i]) for i in range(1 << nv)])
    srs = orc.srs_powers(tau, 1 << nv)
    cm = orc.jac_to_wire(orc.msm(srs, f)).hex()
    assert cm == next(r for r in recs if r[0] == "commitment")[1]


def _mercury_kat(orc, name="mercury_proof.txt"):
    recs = list(golden_lines(name))
    one = lambda key: next(r for r in recs if r[0] == key)[1]
    nv = int(one("numvars"))
    vals = {int(r[1]): r[2] for r in recs if r[0] == "value"}
    pts = {int(r[1]): r[2] for r in recs if r[0] == "point"}
    return dict(
        tau=orc.from_hex_le(orc.FR, one("tau")),
        nv=nv,
        f=np.stack([orc.from_hex_le(orc.FR, vals[i]) for i in range(1 << nv)]),
        u=np.stack([orc.from_hex_le(orc.FR, pts[i]) for i in range(nv)]),
        claim=one("claim"),
        cm=bytes.fromhex(one("commitment")),
        proof=one("proof"),
        probe=one("probe"),
    )


def test_mercury_open_kat(orc):
    """pcs::open over the toy SRS replays the committed full proof at n = 2^4
    byte for byte: the claimed value, the 704 proof bytes (eight points, six
    values) and the terminal transcript sample, which moves with any tag,
    absorb order or squeeze the schedule changes
    (crates/pcs/tests/kats.rs:246-290)."""
    k = _mercury_kat(orc)
    srs = orc.srs_powers(k["tau"], 1 << k["nv"])
    tr = orc.Transcript()
    v, proof = orc.mercury_open(srs, k["f"], k["cm"], k["u"], tr)
    assert orc.to_hex_le(orc.FR, v) == k["claim"]
    assert proof.hex() == k["proof"]
    assert orc.to_hex_le(orc.FR, tr.sample()) == k["probe"]


def test_mercury_open_kat_detects_changes(orc):
    """The replay can fail: one changed witness value or coordinate of u
    moves every proof byte after h."""
    k = _mercury_kat(orc)
    srs = orc.srs_powers(k["tau"], 1 << k["nv"])
    f = k["f"].copy()
    f[3] = orc.from_u64(orc.FR, 12345)
    _, proof = orc.mercury_open(srs, f, k["cm"], k["u"], orc.Transcript())
    assert proof.hex() != k["proof"]
    u = k["u"].copy()
    u[0] = orc.from_u64(orc.FR, 7)
    v, proof = orc.mercury_open(srs, k["f"], k["cm"], u, orc.Transcript())
    assert proof.hex()[:128] != k["proof"][:128] or orc.to_hex_le(orc.FR, v) != k["claim"]
    with pytest.raises(ValueError):
        orc.mercury_open(srs, k["f"][:8], k["cm"], k["u"][:3], orc.Transcript())


def test_mercury_batch_open_kat(orc):
    """batch_open of one column (crates/pcs/src/lib.rs:689-789) replays the
    committed batched proof at n = 2^4: the preamble (the commitment list, u
    then the values, rho) runs before the opening and moves every challenge,
    so this pins its order, tags and encodings (mercury_batch.txt)."""
    k = _mercury_kat(orc, "mercury_batch.txt")
    srs = orc.srs_powers(k["tau"], 1 << k["nv"])
    tr = orc.Transcript()
    vs, proof = orc.mercury_batch_open_stacked(srs, [[k["f"]]], [k["cm"]], k["u"], [], tr)
    assert orc.to_hex_le(orc.FR, vs[0]) == k["claim"]
    assert proof.hex() == k["proof"]
    assert orc.to_hex_le(orc.FR, tr.sample()) == k["probe"]


def test_mercury_batch_open_stacked_properties(orc):
    """Stacks at u || r: each stack's value is sum_j eq(r, j) col_j(u), which
    is the stack's multilinear extension evaluated at u || r (its columns laid
    out slot-major); with one stack of one column and r = [] the batch is a
    bare batch_open. Refusals: an empty batch, a stack with more columns than
    slots."""
    rng = orc.SplitMix(77)
    width, nr = 3, 1
    cols = [np.stack([rng.next_fr() for _ in range(1 << width)]) for _ in range(3)]
    stacks = [[cols[0], cols[1]], [cols[2]]]
    srs = orc.srs_powers(rng.next_fr(), 1 << (width + nr))
    cms = [orc.jac_to_wire(orc.msm(srs[: 1 << (width + nr)], np.concatenate(
        [np.concatenate(s), np.zeros(((2 - len(s)) << width, 4), dtype=np.uint64)]))) for s in stacks]
    u = np.stack([rng.next_fr() for _ in range(width)])
    r = np.stack([rng.next_fr() for _ in range(nr)])
    vs, proof = orc.mercury_batch_open_stacked(srs, stacks, cms, u, r, orc.Transcript())
    for s, v in zip(stacks, vs):
        table = np.concatenate([np.concatenate(s), np.zeros(((2 - len(s)) << width, 4), dtype=np.uint64)])
        want = orc.evaluate(table, np.concatenate([u, r]))
        assert np.array_equal(v, want)
    assert len(proof) == 704
    with pytest.raises(ValueError):
        orc.mercury_batch_open_stacked(srs, [], [], u, r, orc.Transcript())
    with pytest.raises(ValueError):
        orc.mercury_batch_open_stacked(srs, [[cols[0], cols[1], cols[2]]], cms[:1], u, r, orc.Transcript())


@pytest.mark.parametrize("n", [1, 3, 6, 10])
def test_product_claim_and_axpy(orc, n):
    """The claim the full-size shard checks compare with: with one table it
    is the multilinear extension at r (evaluate, pinned by poly_kats); with a
    constant second table it scales by that constant; and it is g_0(0) +
    g_0(1) of the oracle's own product sumcheck (pinned by the brute-force
    definition). axpy accumulates the same fold as combine."""
    rng_seed = 300 + n
    a, b, c = (orc.fill_uniform(rng_seed + i, 1 << n) for i in range(3))
    r = orc.fill_uniform(rng_seed + 9, n)
    assert np.array_equal(orc.product_claim(r, [a]), orc.evaluate(a, r))
    k7 = orc.from_u64(orc.FR, 7)
    seven = np.tile(k7, (1 << n, 1))
    assert np.array_equal(orc.product_claim(r, [a, seven]), orc.mul(orc.FR, orc.evaluate(a, r), k7))
    one = orc.from_u64(orc.FR, 1)
    rounds, _, _ = orc.prove_sumcheck(r, [(one, [0, 1, 2])], [a, b, c], orc.Transcript())
    g0 = rounds[0]
    s = orc.add(orc.FR, g0[0], g0[0])  # g(0) + g(1) = 2 c0 + c1 + c2 + c3 (+ c4)
    for coef in g0[1:]:
        s = orc.add(orc.FR, s, coef)
    assert np.array_equal(orc.product_claim(r, [a, b, c]), s)
    cols = np.stack([a, b, c])
    w = orc.fill_uniform(rng_seed + 11, 3)
    y = orc.fe(1 << n)
    for i in range(3):
        orc.axpy(w[i], cols[i], y)
    assert np.array_equal(y, orc.combine(cols, w, 1)[0])


def test_large_shard_identities_accept_the_oracle_and_reject_changes(orc):
    """The verifier identities the full-size GPU test applies to the large
    shards (tests/test_gpu_mixed_batch.py) hold for the oracle's own shard
    script, and fail when a round coefficient, a final evaluation or an input
    column differs."""
    from types import SimpleNamespace

    from test_gpu_mixed_batch import _large_shard_identities

    nv, k, seed = 6, 3, 8800
    spec = SimpleNamespace(num_vars=nv, columns=k, degree=3, seed=seed)
    cols = np.stack([orc.fill_ctr(seed + c, 1 << nv, "witness") for c in range(k)])
    srs = orc.srs_powers(orc.fill_uniform(5, 1)[0], 1 << nv)
    t = orc.Transcript()
    cms, rounds, fin = orc.shard_script(srs, cols, 3, t)
    proof = SimpleNamespace(commitments=cms, rounds=rounds, final_evals=fin)
    assert _large_shard_identities(orc, spec, proof, t.snapshot()) == []
    bad = rounds.copy()
    bad[2][1] = orc.add(orc.FR, bad[2][1], orc.from_u64(orc.FR, 1))
    assert _large_shard_identities(orc, spec, SimpleNamespace(commitments=cms, rounds=bad, final_evals=fin),
                                   t.snapshot())
    badf = fin.copy()
    badf[1] = orc.add(orc.FR, badf[1], orc.from_u64(orc.FR, 1))
    assert "final evaluation 1" in _large_shard_identities(
        orc, spec, SimpleNamespace(commitments=cms, rounds=rounds, final_evals=badf), t.snapshot())
    other = SimpleNamespace(num_vars=nv, columns=k, degree=3, seed=seed + 1)  # other inputs
    assert "claim" in _large_shard_identities(orc, other, proof, t.snapshot())
