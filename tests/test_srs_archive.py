"""The SRS archive's host half, through the C ABI without a device: the G2
validity check (G2Affine::from_bytes, crates/curve/src/g2.rs:117-199) and the
header rules of Srs::load (crates/srs/src/lib.rs:200-234), with the refusals
of crates/srs/tests/archive.rs:123-192 on archives written here."""
import importlib
import os
import struct

import pytest

import g2_ref as g2

MAGIC = b"APOGESRS"
HEADER = 8 + 4 + 4 + 8 + 128 + 128


@pytest.fixture(scope="module")
def apg():
    return importlib.import_module("apogee-vm_amd")


def test_g2_generator_and_multiples_are_valid(apg):
    assert g2.on_curve(g2.GEN)
    assert apg.g2_wire_valid(g2.wire(g2.GEN))
    assert apg.g2_wire_valid(g2.wire(g2.double(g2.GEN)))
    assert apg.g2_wire_valid(g2.wire(g2.scalar(g2.GEN, 12345)))
    assert apg.g2_wire_valid(bytes(128))  # infinity


def test_g2_refusals(apg):
    w = bytearray(g2.wire(g2.GEN))
    w[0] ^= 1  # off the twist
    assert not apg.g2_wire_valid(bytes(w))
    nc = bytearray(g2.wire(g2.GEN))
    nc[32:64] = (int.from_bytes(nc[32:64], "little") + g2.Q).to_bytes(32, "little")  # x.c1 + q
    assert not apg.g2_wire_valid(bytes(nc))
    # on the twist, outside the order-r subgroup: only r P = O refuses it
    p = g2.point_off_subgroup(20261203)
    assert g2.on_curve(p)
    assert not apg.g2_wire_valid(g2.wire(p))


def archive(power, g1=None, g2_gen=None, g2_tau=None, version=1, count=None):
    count = (1 << power) if count is None else count
    g2_gen = g2.wire(g2.GEN) if g2_gen is None else g2_gen
    g2_tau = g2.wire(g2.double(g2.GEN)) if g2_tau is None else g2_tau
    body = bytes(64 * (1 << power)) if g1 is None else g1
    return MAGIC + struct.pack("<IIQ", version, power, count) + g2_gen + g2_tau + body


def check(apg, tmp_path, name, data):
    path = os.path.join(tmp_path, name)
    with open(path, "wb") as f:
        f.write(data)
    return apg.srs_archive_check(path)


def test_header_accepts_a_wellformed_archive(apg, tmp_path):
    power, g2b = check(apg, tmp_path, "good.srs", archive(3))
    assert power == 3
    assert g2b == g2.wire(g2.GEN) + g2.wire(g2.double(g2.GEN))


def test_header_refusals(apg, tmp_path):
    good = archive(4)
    cases = [
        ("bad-magic", b"X" + good[1:], apg.SrsBadMagic),
        ("bad-count", good[:16] + bytes([5]) + good[17:], apg.SrsBadSection),
        ("bad-power", good[:12] + bytes([200]) + good[13:], apg.SrsBadSection),
        ("short", good[:-1], apg.SrsTruncated),
        ("long", good + b"\0", apg.SrsTruncated),
        ("headless", good[:HEADER - 1], apg.SrsTruncated),
    ]
    for name, data, err in cases:
        with pytest.raises(err):
            check(apg, tmp_path, name, data)
    with pytest.raises(apg.SrsBadVersion) as e:
        check(apg, tmp_path, "bad-version", good[:8] + bytes([9]) + good[9:])
    assert e.value.version == 9
    bad_g2 = bytearray(good)
    bad_g2[24] ^= 1
    with pytest.raises(apg.SrsInvalidPoint) as e:
        check(apg, tmp_path, "bad-g2", bytes(bad_g2))
    assert e.value.index == 16  # the G2 points continue the numbering
    bad_tau = bytearray(good)
    bad_tau[152] ^= 1
    with pytest.raises(apg.SrsInvalidPoint) as e:
        check(apg, tmp_path, "bad-g2-tau", bytes(bad_tau))
    assert e.value.index == 17
    with pytest.raises(apg.SrsIo):
        apg.srs_archive_check(os.path.join(tmp_path, "absent.srs"))


def test_degenerate_header_decodes(apg, tmp_path):
    """tau = 0: g2_tau at infinity decodes (crates/srs/tests/archive.rs:230-262)."""
    power, _ = check(apg, tmp_path, "degenerate.srs", archive(1, g2_tau=bytes(128)))
    assert power == 1
