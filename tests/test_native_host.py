"""CPU-only checks of the native library: it loads without a GPU, exports every
symbol include/apogee_gpu.h declares, and its host-side pieces (the Poseidon2
transcript, the encodings, the interface's error behaviour) match the
reference's vectors. No compute call needs a GPU here."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT, golden_lines
from test_oracle_poly_transcript import replay_cases


@pytest.fixture(scope="session")
def apg():
    import importlib

    return importlib.import_module("apogee-vm_amd")


def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "apogee_gpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(apg_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(apg):
    syms = _declared_symbols()
    assert len(syms) >= 40
    lib = ctypes.CDLL(apg.LIB_PATH)
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_native_transcript_replays_reference_cases(apg, orc):
    class Adapter:
        def __init__(self):
            self.t = apg.Transcript()

        def observe(self, x):
            self.t.observe(x)

        def sample(self):
            return self.t.sample()

        def append_scalars(self, tag, xs):
            self.t.append_scalars(tag, np.asarray(xs).reshape(-1, 4))

        def append_bytes(self, tag, b):
            self.t.append_bytes(tag, b)

        def challenge(self, tag):
            return self.t.challenge_scalar(tag)

    assert replay_cases(orc, Adapter) >= 10


def test_native_permutation_edge_values_match_oracle(apg, orc):
    """The library's permutation keeps its words below 2p between rounds and
    reduces once at the end. States built from 0, 1, p - 1, p - 2 and random
    values, observed and squeezed in turn, agree with the oracle's fully
    reduced permutation (crates/transcript/src/lib.rs:273-292)."""
    pm1 = orc.neg(orc.FR, orc.from_u64(orc.FR, 1))
    pm2 = orc.neg(orc.FR, orc.from_u64(orc.FR, 2))
    rng = orc.SplitMix(77)
    vals = [orc.fe(), orc.from_u64(orc.FR, 1), pm1, pm2] + [rng.next_fr() for _ in range(60)]
    a, b = apg.Transcript(), orc.Transcript()
    for i, v in enumerate(vals):
        a.observe(v)
        b.observe(v)
        if i % 3 == 2:
            assert np.array_equal(a.sample(), b.sample()), i
    for _ in range(40):  # squeezes feed the state back through itself
        assert np.array_equal(a.sample(), b.sample())


def _permute_forms(apg, states):
    """Runs apg_debug_poseidon2_batch in every form the CPU has; returns
    {form: out} (form 2, the mulx/adcx/adox multiply, only where supported)."""
    st = np.ascontiguousarray(states, dtype=np.uint64).reshape(-1, 3, 4)
    outs = {}
    for form in (0, 1, 2):
        out = np.zeros_like(st)
        rc = apg.lib().apg_debug_poseidon2_batch(form, st.ctypes.data, st.shape[0], out.ctypes.data)
        if form == 2 and rc != 0:
            assert "ADX" in apg.lib().apg_last_error().decode()
            continue
        assert rc == 0, (form, apg.lib().apg_last_error())
        outs[form] = out
    return outs


def test_native_permutation_forms_match_reference_vectors(apg, orc):
    """Each host multiply the transcript may dispatch to (the portable 64-bit
    CIOS and the x86-64 mulx/adcx/adox one) gives the reference's
    poseidon2_perm.txt outputs, and agrees with the oracle on states near 0
    and p and on random ones."""
    ins, want = [], []
    for rec in golden_lines("poseidon2_perm.txt"):
        ins.append(np.stack([orc.from_hex_le(orc.FR, h) for h in rec[1:4]]))
        want.append(rec[4:7])
    outs = _permute_forms(apg, np.stack(ins))
    assert len(ins) >= 100 and 0 in outs and 1 in outs
    for form, out in outs.items():
        got = [[orc.to_hex_le(orc.FR, out[i, l]) for l in range(3)] for i in range(len(ins))]
        assert got == want, form
    pm1 = orc.neg(orc.FR, orc.from_u64(orc.FR, 1))
    rng = orc.SplitMix(41)
    edge = [orc.fe(), orc.from_u64(orc.FR, 1), pm1, orc.neg(orc.FR, orc.from_u64(orc.FR, 2))]
    states = [np.stack([edge[(i + l) % 4] for l in range(3)]) for i in range(4)]
    states += [np.stack([rng.next_fr() for _ in range(3)]) for _ in range(200)]
    outs = _permute_forms(apg, np.stack(states))
    for i, s in enumerate(states):
        ref = orc.poseidon2_permute(s)
        for form, out in outs.items():
            assert np.array_equal(out[i], ref), (form, i)
    assert apg.lib().apg_debug_poseidon2_batch(3, None, 0, None) != 0


def test_native_transcript_state_matches_oracle_after_g1(apg, orc):
    recs = list(golden_lines("g1_absorb_kats.txt", {"single"}))
    wires = [bytes.fromhex(r[1]) for r in recs]
    a, b = apg.Transcript(), orc.Transcript()
    a.append_g1_points(3, wires)
    b.append_g1(3, wires)
    assert np.array_equal(a.challenge_scalar(5), b.challenge(5))


def test_encodings(apg, orc):
    rng = orc.SplitMix(3)
    for _ in range(20):
        x = rng.next_fr()
        assert apg.fr_to_bytes(x) == orc.to_bytes(orc.FR, x)
        assert np.array_equal(apg.fr_from_bytes(orc.to_bytes(orc.FR, x)), x)
    p = (0x30644E72E131A029B85045B68181585D2833E84879B9709143E1F593F0000001).to_bytes(32, "little")
    with pytest.raises(apg.ApgError):
        apg.fr_from_bytes(p)
    assert np.array_equal(apg.fr_from_u64(12345), orc.from_u64(orc.FR, 12345))
    for rec in golden_lines("g1_kats_subset.txt", {"g1_ops"}):
        w = bytes.fromhex(rec[1])
        m = apg.g1_from_wire(w)
        assert apg.g1_to_wire(m) == w
        assert np.array_equal(np.frombuffer(m, dtype=np.uint64), orc.affine_from_wire(w))
    for rec in golden_lines("g1_kats_subset.txt", {"g1_reject"}):
        with pytest.raises(apg.ApgError):
            apg.g1_from_wire(bytes.fromhex(rec[1]))
    assert apg.g1_to_wire(bytes(64)) == bytes(64)


def test_interface_errors(apg):
    one = apg.fr_from_u64(1)
    with pytest.raises(apg.GateError):
        apg.Gate([], [])
    with pytest.raises(apg.GateError):
        apg.Gate([1, 1], [])
    with pytest.raises(apg.GateError):
        apg.Gate([1, 2], [apg.GateTerm(one, 0, 2)])
    for nv in (0, 1, 3, 55, 56):
        with pytest.raises(apg.UnsupportedNumVars):
            apg.check_num_vars(nv)
    assert apg.check_num_vars(4) == 16
    with pytest.raises(apg.LengthMismatch):
        apg.msm(np.zeros((3, 8), dtype=np.uint64), np.zeros((2, 4), dtype=np.uint64))
    assert apg.Gate.product(2).degree() == 3
    assert apg.Gate.product(3).degree() == 4
    assert apg.Gate([0], [apg.GateTerm(one, 0)]).degree() == 3


def test_abi_refusals_without_a_device(apg):
    """Argument errors come back as APG_ERR_ARG before anything touches a
    device: a shard of degree 5 (a gate term holds three factors, so a
    shard's summand is of degree 3 or 4), null handles, and engine ids never
    issued."""
    L = apg.lib()
    err = lambda: L.apg_last_error().decode()
    buf = np.zeros(16, dtype=np.uint64)
    for deg in (2, 5, 6):
        rc = L.apg_prove_shard_dev(None, None, None, 4, 2, deg, None, buf.ctypes.data, buf.ctypes.data,
                                   buf.ctypes.data)
        assert rc == -2 and "degree" in err()
    for deg in (3, 4):  # a valid degree gets as far as the null context
        rc = L.apg_prove_shard_dev(None, None, None, 4, 2, deg, None, buf.ctypes.data, buf.ctypes.data,
                                   buf.ctypes.data)
        assert rc == -2 and "null handle" in err()
    jid = ctypes.c_uint64()
    assert L.apg_engine_submit_shard(None, None, None, 4, 2, 5, None, ctypes.byref(jid)) == -2
    assert "degree" in err()
    assert L.apg_engine_submit_shard(None, None, None, 4, 2, 3, None, ctypes.byref(jid)) == -2
    assert L.apg_ctx_synchronize(None) == -2
    assert L.apg_engine_poll(None, 1) == -2
    assert L.apg_engine_wait(None, 1, None, None, None, None, None) == -2
    assert L.apg_engine_stats(None, None) == -2
    assert L.apg_engine_set_policy(None, 0, 0, ctypes.c_double(0.0)) == -2


def _mul32(apg, which, a, b):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    b = np.ascontiguousarray(b, dtype=np.uint64)
    out = np.zeros_like(a)
    rc = apg.lib().apg_debug_mul32_batch(which, a.ctypes.data, b.ctypes.data, a.shape[0], out.ctypes.data)
    assert rc == 0
    return out


@pytest.mark.parametrize("which,modulus", [
    (0, 0x30644E72E131A029B85045B68181585D2833E84879B9709143E1F593F0000001),
    (1, 0x30644E72E131A029B85045B68181585D97816A916871CA8D3C208C16D87CFD47),
])
def test_device_montgomery_mul_matches_oracle(apg, orc, which, modulus):
    """The kernels' 32-bit-word no-carry CIOS (compiled for the host) equals
    the reference's u64 CIOS on random inputs and on the edge grid where a
    carry bound would break first (m - 1, m - 2^k, 2^k, R, R^2, ...)."""
    rng = np.random.default_rng(1000 + which)
    n = 100_000
    vals = [int.from_bytes(rng.bytes(32), "little") % modulus for _ in range(n)]
    edges = [0, 1, 2, modulus - 1, modulus - 2, (modulus - 1) // 2, (modulus + 1) // 2]
    edges += [(1 << k) % modulus for k in range(0, 256, 7)]
    edges += [modulus - (1 << k) for k in range(0, 250, 5)]
    edges += [(1 << 256) % modulus, (1 << 512) % modulus, modulus - ((1 << 256) % modulus)]
    pairs_a = vals + [e for e in edges for _ in edges]
    pairs_b = vals[::-1] + [f for _ in edges for f in edges]

    def limbs(xs):
        return np.array([[(x >> (64 * i)) & ((1 << 64) - 1) for i in range(4)] for x in xs], dtype=np.uint64)

    a, b = limbs(pairs_a), limbs(pairs_b)
    got = _mul32(apg, which, a, b)
    want = np.zeros_like(a)
    L = orc.lib()
    for i in range(a.shape[0]):
        L.orc_mul(which, orc.ptr(a[i]), orc.ptr(b[i]), orc.ptr(want[i]))
    assert np.array_equal(got, want)


FQ_MODULUS = 0x30644E72E131A029B85045B68181585D97816A916871CA8D3C208C16D87CFD47


def _limbs(xs):
    return np.array([[(x >> (64 * i)) & ((1 << 64) - 1) for i in range(4)] for x in xs], dtype=np.uint64)


def fq29_cases(orc, n=40_000):
    """Operand pairs for the lazy 29-bit Fq and the oracle's results for ops
    0..5: random values, the edge grid (0, 1, p - 1, 2^k, p - 2^k, R mod p),
    and non-canonical inputs x + p (below 2p), which the lazy form accepts."""
    p = FQ_MODULUS
    rng = np.random.default_rng(2029)
    vals = [int.from_bytes(rng.bytes(32), "little") % p for _ in range(n)]
    edges = [0, 1, 2, p - 1, p - 2, (p - 1) // 2, (p + 1) // 2, (1 << 256) % p, p - ((1 << 256) % p)]
    edges += [(1 << k) % p for k in range(0, 254, 9)] + [p - (1 << k) for k in range(0, 250, 11)]
    xa = vals + [e for e in edges for _ in edges]
    xb = vals[::-1] + [f for _ in edges for f in edges]
    ra = [x + p if i % 3 == 0 else x for i, x in enumerate(xa)]
    rb = [x + p if i % 3 == 1 else x for i, x in enumerate(xb)]
    assert max(ra + rb) < 2 * p
    a, b = _limbs(xa), _limbs(xb)
    L = orc.lib()
    want = {op: np.zeros_like(a) for op in range(6)}
    s, d, q = (np.zeros(4, dtype=np.uint64) for _ in range(3))
    for i in range(a.shape[0]):
        pa, pb = orc.ptr(a[i]), orc.ptr(b[i])
        L.orc_mul(1, pa, pb, orc.ptr(want[0][i]))
        L.orc_add(1, pa, pb, orc.ptr(want[1][i]))
        L.orc_sub(1, pa, pb, orc.ptr(want[2][i]))
        L.orc_mul(1, orc.ptr(want[1][i]), orc.ptr(want[2][i]), orc.ptr(d))
        L.orc_mul(1, pa, pa, orc.ptr(q))
        L.orc_add(1, orc.ptr(d), orc.ptr(q), orc.ptr(s))
        L.orc_sub(1, orc.ptr(s), pb, orc.ptr(want[3][i]))
        L.orc_mul(1, pa, pa, orc.ptr(want[4][i]))
        L.orc_mul(1, pa, pb, orc.ptr(d))
        L.orc_add(1, orc.ptr(d), orc.ptr(want[4][i]), orc.ptr(want[5][i]))
    return _limbs(ra), _limbs(rb), want


def fq29_run(apg, ctx_handle, op, a, b):
    out = np.zeros_like(a)
    rc = apg.lib().apg_debug_fq29_batch(ctx_handle, op, a.ctypes.data, b.ctypes.data, a.shape[0], out.ctypes.data)
    assert rc == 0, apg.lib().apg_last_error()
    return out


def test_fq29_lazy_arithmetic_matches_oracle(apg, orc):
    """The MSM accumulation's Fq in nine 29-bit words with lazy reduction
    (results kept in [0, 2p)), compiled for the host, against the oracle's
    u64 Montgomery Fq: multiply, add, subtract, a chain through the whole
    lazy range, the squaring and two products under one reduction, with
    non-canonical inputs and the edge grid."""
    a, b, want = fq29_cases(orc)
    for op in range(6):
        assert np.array_equal(fq29_run(apg, None, op, a, b), want[op]), op
    assert apg.lib().apg_debug_fq29_batch(None, 6, a.ctypes.data, b.ctypes.data, 1, a.ctypes.data) != 0
    off = np.zeros(2, dtype=np.uint64)
    assert apg.lib().apg_debug_accumulate29(None, 3, a.ctypes.data, a.ctypes.data, off.ctypes.data, 1,
                                            a.ctypes.data) != 0


def accumulate29_cases(orc):
    """Chains of signed affine points and the oracle's sums (Jacobian chain of
    affine additions, crates/curve/src/g1.rs:203-243). Points are small
    multiples of G with random signs, so the accumulator hits +-P often
    (doubling and identity), plus random points with infinities mixed in."""
    L = orc.lib()
    rng = np.random.default_rng(29)
    small = np.zeros((4, 4), dtype=np.uint64)
    for i in range(4):
        L.orc_from_u64(0, i + 1, orc.ptr(small[i]))
    big = np.zeros((64, 4), dtype=np.uint64)
    for i in range(64):
        L.orc_from_u64(0, int(rng.integers(1, 2**63)), orc.ptr(big[i]))
    pts_small = np.zeros((4, 8), dtype=np.uint64)
    pts_big = np.zeros((64, 8), dtype=np.uint64)
    L.orc_g1_gen_mul_batch(orc.ptr(small), 4, 1, orc.ptr(pts_small))
    L.orc_g1_gen_mul_batch(orc.ptr(big), 64, 1, orc.ptr(pts_big))
    inf = np.zeros(8, dtype=np.uint64)

    def oracle_sum(pts, neg):
        acc = np.zeros(12, dtype=np.uint64)
        L.orc_g1_from_affine(orc.ptr(inf), orc.ptr(acc))
        for p, s in zip(pts, neg):
            q = p.copy()
            if s and q.any():
                L.orc_neg(1, orc.ptr(p[4:]), orc.ptr(q[4:]))
            nxt = np.zeros(12, dtype=np.uint64)
            L.orc_g1_add_affine(orc.ptr(acc), orc.ptr(q), orc.ptr(nxt))
            acc = nxt
        out = np.zeros(8, dtype=np.uint64)
        L.orc_g1_to_affine(orc.ptr(acc), orc.ptr(out))
        return out

    P, Q = pts_small[0], pts_small[1]
    cases = [
        ([], []),
        ([P], [0]),
        ([P, P], [0, 0]),            # doubling
        ([P, P], [1, 1]),            # doubling of -P
        ([P, P], [0, 1]),            # identity
        ([P, P, Q], [0, 1, 0]),      # identity, then restart
        ([P, P, P, P], [0, 1, 0, 0]),
        ([P, inf, Q, inf], [0, 0, 1, 1]),
        ([inf, inf], [0, 1]),
    ]
    for _ in range(40):
        m = int(rng.integers(1, 40))
        idx = rng.integers(0, 4, size=m)
        cases.append(([pts_small[i] for i in idx], list(rng.integers(0, 2, size=m))))
    for _ in range(8):
        m = int(rng.integers(1, 64))
        idx = rng.integers(0, 64, size=m)
        pts = [pts_big[i] for i in idx]
        for j in rng.integers(0, m, size=m // 8):
            pts[j] = inf
        cases.append((pts, list(rng.integers(0, 2, size=m))))
    flat = [p for pts, _ in cases for p in pts]
    points = np.ascontiguousarray(np.array(flat, dtype=np.uint64).reshape(-1, 8))
    neg = np.ascontiguousarray(np.array([s for _, ng in cases for s in ng], dtype=np.uint8))
    offsets = np.zeros(len(cases) + 1, dtype=np.uint64)
    offsets[1:] = np.cumsum([len(pts) for pts, _ in cases])
    want = np.array([oracle_sum(pts, ng) for pts, ng in cases], dtype=np.uint64).reshape(-1, 8)
    return points, neg, offsets, want


def accumulate29_run(apg, ctx_handle, form, points, neg, offsets):
    n = offsets.shape[0] - 1
    out = np.zeros((n, 8), dtype=np.uint64)
    rc = apg.lib().apg_debug_accumulate29(ctx_handle, form, points.ctypes.data, neg.ctypes.data,
                                          offsets.ctypes.data, n, out.ctypes.data)
    assert rc == 0, apg.lib().apg_last_error()
    return out


@pytest.mark.parametrize("form", [0, 1, 2])
def test_accumulate29_matches_oracle_chain(apg, orc, form):
    """The MSM's point sums, compiled for the host, against the oracle's
    chains: form 0 is the bucket loop in the lazy 29-bit arithmetic (mixed
    XYZZ adds, affine doubling when the point equals the accumulator, identity
    when it is its negation, infinity skipped), form 1 the same in the 32-bit
    form, form 2 the bucket reduction's general XYZZ add in the lazy form, on
    the points rescaled to Z != 1 (so equal points with different Z reach its
    doubling branch). The empty chain is the identity (all-zero)."""
    points, neg, offsets, want = accumulate29_cases(orc)
    got = accumulate29_run(apg, None, form, points, neg, offsets)
    for c in range(want.shape[0]):
        assert np.array_equal(got[c], want[c]), c
    assert not got[0].any()

@pytest.mark.parametrize("kind", ["uniform", "witness"])
def test_synthetic_inputs_match_oracle(apg, orc, kind):
    """The benchmark's counter-mode ChaCha20 draws (host path of the library)
    equal the oracle's restatement; the witness mix has its stated shares."""
    a = apg.synthetic(12345, 20000, kind)
    b = orc.fill_ctr(12345, 20000, kind)
    assert np.array_equal(a, b)
    if kind == "witness":
        zero = (a == 0).all(1).mean()
        one = (a == orc.from_u64(orc.FR, 1)).all(1).mean()
        assert abs(zero - 0.2) < 0.02 and abs(one - 0.1) < 0.02


@pytest.mark.parametrize("k,nv", [(1, 0), (2, 3), (3, 7)])
def test_witness_digest_matches_oracle(apg, orc, k, nv):
    """sumcheck::witness_digest (host path of the library) equals the
    oracle's sponge, and a one-cell change moves it."""
    cols = [orc.fill_uniform(500 + c, 1 << nv) for c in range(k)]
    d = apg.witness_digest(cols)
    assert np.array_equal(d, orc.witness_digest(cols))
    cols[-1][-1] = orc.from_u64(orc.FR, 7)
    assert not np.array_equal(apg.witness_digest(cols), d)
    t = apg.Transcript()
    apg.absorb_witness_digest(t, d)
    ot = orc.Transcript()
    ot.append_scalars(orc.TAG_WITNESS_DIGEST, d.reshape(1, 4))
    assert t.state()[:96] == ot.snapshot()[:96]


def test_witness_digest_rejects_no_columns(apg):
    with pytest.raises(ValueError):
        apg.witness_digest([])


def test_transcript_state_round_trip(apg, orc):
    """apg_transcript_set_state inverts apg_transcript_state: a transcript
    moved across mid-message (one scalar in the rate buffer, a squeeze
    pending) continues exactly as the original and as the oracle; malformed
    states are refused."""
    rng = orc.SplitMix(31)
    a, o = apg.Transcript(), orc.Transcript()
    for step in range(5):
        xs = [rng.next_fr() for _ in range(step + 1)]
        a.append_scalars(6, xs)
        o.append_scalars(6, xs)
        a.observe(xs[0])
        o.observe(xs[0])
        b = apg.Transcript()
        b.set_state(a.state())
        assert b.state() == a.state()
        for t in (a, b):
            t.append_scalars(4, xs[:1])
        o.append_scalars(4, xs[:1])
        ca, cb, co = a.challenge_scalar(5), b.challenge_scalar(5), o.challenge(5)
        assert np.array_equal(ca, cb) and np.array_equal(ca, co)
    bad = bytearray(a.state())
    bad[224] = 3
    with pytest.raises(apg.ApgError):
        apg.Transcript().set_state(bytes(bad))
    bad = bytearray(a.state())
    bad[0:32] = b"\xff" * 32
    with pytest.raises(apg.ApgError):
        apg.Transcript().set_state(bytes(bad))
