"""GPU parity for SRS ingestion through the C ABI: G1 points decoded and
validated on the device against the oracle's decoder, the archive round trip
and its pinned layout, the refusals of crates/srs/tests/archive.rs:123-192,
and batch_to_affine on the device against the oracle."""
import os
import struct

import numpy as np
import pytest

import g2_ref as g2

pytestmark = pytest.mark.gpu

MAGIC = b"APOGESRS"
HEADER = 8 + 4 + 4 + 8 + 128 + 128
G2B = g2.wire(g2.GEN) + g2.wire(g2.double(g2.GEN))


@pytest.fixture(scope="module")
def apg():
    import importlib

    m = importlib.import_module("apogee-vm_amd")
    m.default_context()
    return m


def _powers(orc, n, seed=7):
    tau = orc.fill_uniform(seed, 1)[0]
    return orc.srs_powers(tau, n)


def _wire(orc, pts):
    return b"".join(orc.affine_to_wire(p) for p in pts)


def test_from_wire_matches_oracle(apg, orc):
    pts = _powers(orc, 300)
    pts[17] = 0  # infinity
    srs = apg.Srs.from_wire(_wire(orc, pts))
    print("srs from_wire n=300", flush=True)
    assert np.array_equal(srs.g1(), pts)


def test_from_wire_refusals(apg, orc):
    pts = _powers(orc, 64)
    w = bytearray(_wire(orc, pts))
    w[9 * 64] ^= 1  # off the curve
    w[7 * 64:7 * 64 + 32] = b"\xff" * 32  # non-canonical x
    with pytest.raises(apg.SrsInvalidPoint) as e:
        apg.Srs.from_wire(bytes(w))
    assert e.value.index == 7  # the lowest invalid index
    # the oracle's decoder refuses the same two points and no other
    bad = [i for i in range(64) if orc.affine_from_wire(bytes(w[64 * i:64 * i + 64])) is None]
    assert bad == [7, 9]


@pytest.mark.parametrize("power", [0, 12, 21])
def test_save_load_round_trip(apg, orc, tmp_path, power):
    tau = orc.fill_uniform(11, 1)[0]
    src = apg.Srs.generate(tau, 1 << power)
    path = os.path.join(tmp_path, f"rt-{power}.srs")
    src.save(path, G2B)
    back, g2b = apg.Srs.load(path)
    print(f"srs round trip power={power}", flush=True)
    assert len(back) == 1 << power and g2b == G2B
    assert np.array_equal(back.g1(), src.g1())
    assert os.path.getsize(path) == HEADER + 64 * (1 << power)
    # save . load . save is save
    again = os.path.join(tmp_path, f"rt2-{power}.srs")
    back.save(again, g2b)
    assert open(path, "rb").read() == open(again, "rb").read()


def test_header_layout_is_pinned(apg, orc, tmp_path):
    tau = orc.fill_uniform(12, 1)[0]
    srs = apg.Srs.generate(tau, 1 << 12)
    path = os.path.join(tmp_path, "layout.srs")
    srs.save(path, G2B)
    b = open(path, "rb").read()
    assert b[:8] == MAGIC
    assert struct.unpack("<IIQ", b[8:24]) == (1, 12, 4096)
    assert b[24:280] == G2B
    gen = (1).to_bytes(32, "little") + (2).to_bytes(32, "little")  # G1 generator (1, 2)
    assert b[280:344] == gen
    assert b[344:408] == orc.affine_to_wire(srs.g1(1, 1)[0])


def test_tampered_archives(apg, orc, tmp_path):
    tau = orc.fill_uniform(13, 1)[0]
    srs = apg.Srs.generate(tau, 1 << 12)
    good_path = os.path.join(tmp_path, "good.srs")
    srs.save(good_path, G2B)
    good = open(good_path, "rb").read()

    def load(name, data):
        p = os.path.join(tmp_path, name)
        with open(p, "wb") as f:
            f.write(data)
        return apg.Srs.load(p)

    b = bytearray(good)
    b[HEADER + 5 * 64] ^= 1
    with pytest.raises(apg.SrsInvalidPoint) as e:
        load("bad-point.srs", bytes(b))
    assert e.value.index == 5
    b = bytearray(good)
    b[HEADER + 7 * 64:HEADER + 7 * 64 + 32] = b"\xff" * 32
    with pytest.raises(apg.SrsInvalidPoint) as e:
        load("noncanonical.srs", bytes(b))
    assert e.value.index == 7
    b = bytearray(good)
    b[24] ^= 1
    with pytest.raises(apg.SrsInvalidPoint) as e:
        load("bad-g2.srs", bytes(b))
    assert e.value.index == 4096
    with pytest.raises(apg.SrsBadMagic):
        load("bad-magic.srs", b"X" + good[1:])
    with pytest.raises(apg.SrsBadVersion):
        load("bad-version.srs", good[:8] + bytes([9]) + good[9:])
    with pytest.raises(apg.SrsBadSection):
        load("bad-count.srs", good[:16] + bytes([5]) + good[17:])
    with pytest.raises(apg.SrsTruncated):
        load("short.srs", good[:-1])
    with pytest.raises(apg.SrsTruncated):
        load("long.srs", good + b"\0")
    with pytest.raises(apg.SrsIo):
        apg.Srs.load(os.path.join(tmp_path, "absent.srs"))
    # a real point in the wrong place still decodes
    b = bytearray(good)
    b[HEADER + 5 * 64:HEADER + 6 * 64] = good[HEADER + 6 * 64:HEADER + 7 * 64]
    swapped, _ = load("swapped.srs", bytes(b))
    assert np.array_equal(swapped.g1(5, 1), srs.g1(6, 1))


def test_degenerate_archive_loads(apg, tmp_path):
    """tau = 0 (crates/srs/tests/archive.rs:230-262): infinity points decode."""
    data = (MAGIC + struct.pack("<IIQ", 1, 1, 2) + g2.wire(g2.GEN) + bytes(128)
            + (1).to_bytes(32, "little") + (2).to_bytes(32, "little") + bytes(64))
    p = os.path.join(tmp_path, "degenerate.srs")
    with open(p, "wb") as f:
        f.write(data)
    srs, g2b = apg.Srs.load(p)
    pts = srs.g1()
    assert pts[0].any() and not pts[1].any()
    assert g2b[128:] == bytes(128)


def test_batch_to_affine_matches_oracle(apg, orc):
    g = orc.jac_from_affine(orc.generator_affine())
    jac = []
    for i in range(257):
        s = orc.fill_uniform(100 + i, 1)[0]
        p = orc.jac_mul(g, s)
        jac.append(p)
    jac.append(np.zeros(12, dtype=np.uint64))  # z = 0: the identity
    jac = np.stack(jac)
    got = apg.batch_to_affine(jac)
    want = np.stack([orc.jac_to_affine(p) for p in jac[:-1]] + [np.zeros(8, dtype=np.uint64)])
    print("batch_to_affine n=258", flush=True)
    assert np.array_equal(got, want)
