"""Oracle Fr/Fq arithmetic against the reference's committed vectors
(crates/field/tests/vectors/fr_kats.txt, crates/curve/tests/vectors/fq_kats.txt)."""
from conftest import golden_lines
import numpy as np


def _fe(orc, which, h):
    x = orc.from_hex_le(which, h)
    assert x is not None, h
    return x


def test_fr_kats(orc):
    n = 0
    for rec in golden_lines("fr_kats.txt"):
        op = rec[0]
        a = _fe(orc, orc.FR, rec[1])
        if op in ("add", "sub", "mul"):
            b = _fe(orc, orc.FR, rec[2])
            got = getattr(orc, op)(orc.FR, a, b)
            assert orc.to_hex_le(orc.FR, got) == rec[3], rec
        elif op == "square":
            assert orc.to_hex_le(orc.FR, orc.square(orc.FR, a)) == rec[2]
        elif op == "inv":
            inv = orc.inverse(orc.FR, a)
            if rec[2] == "none":
                assert inv is None
            else:
                assert orc.to_hex_le(orc.FR, inv) == rec[2]
        elif op == "pow":
            e = np.frombuffer(bytes.fromhex(rec[2]), dtype=np.uint64).copy()
            assert orc.to_hex_le(orc.FR, orc.pow_(orc.FR, a, e)) == rec[3]
        else:
            raise AssertionError(op)
        n += 1
    assert n == 536


def test_fr_corrupted_vector_fails(orc):
    rec = next(r for r in golden_lines("fr_kats.txt", {"mul"}) if r[1] != "0" * 64 and r[2] != "0" * 64)
    a, b = _fe(orc, orc.FR, rec[1]), _fe(orc, orc.FR, rec[2])
    bad = ("1" if rec[3][0] == "0" else "0") + rec[3][1:]
    assert orc.to_hex_le(orc.FR, orc.mul(orc.FR, a, b)) != bad


def test_fq_kats(orc):
    n = 0
    for rec in golden_lines("fq_kats_subset.txt", {"fq_ops", "fq_pow", "fq_bytes", "fq_u64"}):
        if rec[0] == "fq_ops":
            a, b = _fe(orc, orc.FQ, rec[1]), _fe(orc, orc.FQ, rec[2])
            assert orc.to_hex_le(orc.FQ, orc.add(orc.FQ, a, b)) == rec[3]
            assert orc.to_hex_le(orc.FQ, orc.sub(orc.FQ, a, b)) == rec[4]
            assert orc.to_hex_le(orc.FQ, orc.mul(orc.FQ, a, b)) == rec[5]
            assert orc.to_hex_le(orc.FQ, orc.square(orc.FQ, a)) == rec[6]
            inv = orc.inverse(orc.FQ, a)
            assert (inv is None) == (rec[7] == "none")
            if inv is not None:
                assert orc.to_hex_le(orc.FQ, inv) == rec[7]
        elif rec[0] == "fq_pow":
            a = _fe(orc, orc.FQ, rec[1])
            e = np.frombuffer(bytes.fromhex(rec[2]), dtype=np.uint64).copy()
            assert orc.to_hex_le(orc.FQ, orc.pow_(orc.FQ, a, e)) == rec[3]
        elif rec[0] == "fq_bytes":
            ok = orc.from_hex_le(orc.FQ, rec[1]) is not None
            assert ok == (rec[2] == "ok"), rec
        elif rec[0] == "fq_u64":
            assert orc.to_hex_le(orc.FQ, orc.from_u64(orc.FQ, int(rec[1]))) == rec[2]
        n += 1
    assert n > 300


def test_montgomery_constants(orc):
    # R mod p is ONE; R^2 converts; -1 = p - 1 (crates/field/src/lib.rs:22-27)
    one = orc.from_u64(orc.FR, 1)
    assert orc.to_bytes(orc.FR, one) == (1).to_bytes(32, "little")
    p = 0x30644E72E131A029B85045B68181585D2833E84879B9709143E1F593F0000001
    m1 = orc.sub(orc.FR, orc.fe(), one)
    assert int.from_bytes(orc.to_bytes(orc.FR, m1), "little") == p - 1
    # Python big-int multiplication as the closed form
    rng = np.random.default_rng(5)
    for _ in range(50):
        a = int.from_bytes(rng.bytes(32), "little") % p
        b = int.from_bytes(rng.bytes(32), "little") % p
        fa = orc.from_bytes(orc.FR, a.to_bytes(32, "little"))
        fb = orc.from_bytes(orc.FR, b.to_bytes(32, "little"))
        assert int.from_bytes(orc.to_bytes(orc.FR, orc.mul(orc.FR, fa, fb)), "little") == a * b % p
