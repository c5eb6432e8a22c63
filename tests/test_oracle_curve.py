"""Oracle G1 and MSM against the reference's committed vectors
(crates/curve/tests/vectors/g1_kats.txt, msm_kats.txt) plus the closed forms
crates/curve/tests/msm.rs checks (small path == general path, thread-count
independence, brute-force sum on tiny inputs)."""
import numpy as np
import pytest

from conftest import golden_lines

INF = "0" * 128


def _pt(orc, h):
    p = orc.affine_from_wire(bytes.fromhex(h))
    assert p is not None, h
    return p


def _hx(orc, aff):
    return orc.affine_to_wire(aff).hex()


def test_g1_kats(orc):
    counts = {}
    for rec in golden_lines("g1_kats_subset.txt"):
        kind = rec[0]
        counts[kind] = counts.get(kind, 0) + 1
        if kind == "g1_generator":
            assert _hx(orc, orc.generator_affine()) == rec[1]
        elif kind == "g1_ops":
            p, q = _pt(orc, rec[1]), _pt(orc, rec[2])
            k = orc.from_hex_le(orc.FR, rec[3])
            pj, qj = orc.jac_from_affine(p), orc.jac_from_affine(q)
            assert _hx(orc, orc.jac_to_affine(orc.jac_add(pj, qj))) == rec[4]
            assert _hx(orc, orc.jac_to_affine(orc.jac_add_affine(pj, q))) == rec[4]
            assert _hx(orc, orc.jac_to_affine(orc.jac_double(pj))) == rec[5]
            assert _hx(orc, orc.jac_to_affine(orc.jac_mul(pj, k))) == rec[6]
            negp = p.copy()
            negp[4:] = orc.neg(orc.FQ, p[4:])
            assert _hx(orc, negp) == rec[7]
        elif kind == "g1_add_edge":
            p, q = _pt(orc, rec[2]), _pt(orc, rec[3])
            pj, qj = orc.jac_from_affine(p), orc.jac_from_affine(q)
            assert _hx(orc, orc.jac_to_affine(orc.jac_add(pj, qj))) == rec[4], rec[1]
            assert _hx(orc, orc.jac_to_affine(orc.jac_add_affine(pj, q))) == rec[4], rec[1]
        elif kind == "g1_madd_edge":
            p, two_p, four_p, neg_two_p = (_pt(orc, h) for h in rec[1:5])
            acc = orc.jac_double(orc.jac_from_affine(p))  # Z != 1
            assert _hx(orc, orc.jac_to_affine(acc)) == rec[2]
            assert _hx(orc, orc.jac_to_affine(orc.jac_add_affine(acc, two_p))) == rec[3]
            assert _hx(orc, orc.jac_to_affine(orc.jac_add_affine(acc, neg_two_p))) == INF
        elif kind == "g1_scalar_edge":
            p = _pt(orc, rec[2])
            kb = bytes.fromhex(rec[3])
            got = orc.jac_to_affine(orc.jac_mul_le_bytes(orc.jac_from_affine(p), kb))
            assert _hx(orc, got) == rec[4], rec[1]
        elif kind == "g1_reject":
            assert orc.affine_from_wire(bytes.fromhex(rec[1])) is None, rec[2]
    assert counts["g1_ops"] == 120 and counts["g1_add_edge"] == 6 and counts["g1_reject"] == 10


def msm_case(orc, pattern, n, seed):
    """crates/curve/tests/msm.rs:41-68: n bases k_i*G, then n scalars."""
    rng = orc.SplitMix(seed)
    ks = np.stack([rng.next_fr() for _ in range(n)]) if n else orc.fe(0)
    bases = orc.gen_mul_batch(ks) if n else np.zeros((0, 8), dtype=np.uint64)
    if pattern == "infinity":
        bases[0::3] = 0
    draws = np.stack([rng.next_fr() for _ in range(n)]) if n else orc.fe(0)
    if pattern == "zeros":
        scalars = orc.fe(n)
    elif pattern == "identical":
        scalars = np.repeat(draws[:1], n, axis=0) if n else orc.fe(0)
    else:
        scalars = draws
    return np.ascontiguousarray(bases), np.ascontiguousarray(scalars)


def test_msm_kats(orc):
    checked = 0
    for rec in golden_lines("msm_kats.txt"):
        _, pattern, n, seed, want = rec
        bases, scalars = msm_case(orc, pattern, int(n), int(seed))
        got = orc.jac_to_wire(orc.msm(bases, scalars, threads=8, workers=8)).hex()
        assert got == want, (pattern, n)
        checked += 1
    assert checked == 13


def test_msm_corrupted_vector_fails(orc):
    rec = next(r for r in golden_lines("msm_kats.txt") if r[1] == "random" and r[2] == "100")
    bases, scalars = msm_case(orc, "random", 100, int(rec[3]))
    got = orc.jac_to_wire(orc.msm(bases, scalars)).hex()
    bad = ("1" if rec[4][0] == "0" else "0") + rec[4][1:]
    assert got == rec[4] and got != bad


@pytest.mark.parametrize("n", [0, 1, 2, 3, 5, 31, 32, 33, 100])
def test_msm_brute_force(orc, n):
    rng = orc.SplitMix(1000 + n)
    bases, scalars = msm_case(orc, "random", n, 77 + n)
    acc = orc.jac_from_affine(np.zeros(8, dtype=np.uint64))
    for i in range(n):
        acc = orc.jac_add(acc, orc.jac_mul(orc.jac_from_affine(bases[i]), scalars[i]))
    assert orc.jac_to_wire(orc.msm(bases, scalars)) == orc.jac_to_wire(acc)
    del rng


@pytest.mark.parametrize("n", [0, 1, 33, 1000, 4096])
def test_msm_small_path_and_threads(orc, n):
    rng = np.random.default_rng(n)
    bases, _ = msm_case(orc, "random", n, 4242 + n)
    small = rng.integers(0, 2**32, size=n, dtype=np.uint64).astype(np.uint32)
    small[: n // 4] = 0xFFFFFFFF
    lifted = np.stack([orc.from_u64(orc.FR, int(s)) for s in small]) if n else orc.fe(0)
    a = orc.jac_to_wire(orc.msm_small_u32(bases, small))
    b = orc.jac_to_wire(orc.msm(bases, lifted))
    assert a == b
    for threads in (1, 3, 17, 128):
        assert orc.jac_to_wire(orc.msm(bases, lifted, threads=threads, workers=4)) == b


def test_window_rule(orc):
    # crates/curve/src/msm.rs:112-128
    assert orc.window_bits(31) == 3
    assert orc.window_bits(1 << 16) == 13
    assert orc.window_bits(1 << 20) == 15
    assert orc.window_bits(1 << 22) == 17
    assert orc.window_bits(33) == 6 * 69 // 100 + 2


@pytest.mark.parametrize("w", [3, 13, 15, 17])
def test_recode_digits_and_nonzero_count(orc, w):
    """The reference's signed-digit recoding (crates/curve/src/msm.rs:180-198),
    which the survey's algorithmic MSM work counts: digits reconstruct the
    canonical scalar (sum d_j 2^(w j)), lie in [-2^(w-1), 2^(w-1)], and the
    batch count equals the per-scalar nonzero digits; 0 has none, 1 one,
    2^w - 1 two (-1, then the carry)."""
    R = 2**256
    p = 0x30644E72E131A029B85045B68181585D2833E84879B9709143E1F593F0000001
    xs = [orc.fe(), orc.from_u64(orc.FR, 1), orc.from_u64(orc.FR, (1 << w) - 1)]
    xs += list(orc.fill_uniform(41 + w, 40)) + list(orc.fill_witness(43 + w, 40))
    total = 0
    for x in xs:
        d = orc.msm_recode(x, w)
        half = 1 << (w - 1)
        assert all(-half <= int(v) <= half for v in d)
        val = sum(int(v) << (w * j) for j, v in enumerate(d))
        mont = sum(int(x[i]) << (64 * i) for i in range(4))
        assert val == mont * pow(R, -1, p) % p
        total += int(np.count_nonzero(d))
    assert [int(np.count_nonzero(orc.msm_recode(x, w))) for x in xs[:3]] == [0, 1, 2]
    assert orc.msm_nonzero_digits(np.stack(xs), w) == total
