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
