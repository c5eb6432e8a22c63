"""pcs::open on the GPU (crates/pcs/src/lib.rs:313-492) against the committed
full proof at n = 2^4 (crates/pcs/tests/vectors/mercury_proof.txt) and against
the oracle's restatement (pinned by the same proof in
test_oracle_sumcheck.py) on random instances up to 2^16: the claimed value, the
704 proof bytes and the transcript state after the opening agree byte for
byte."""
import numpy as np
import pytest

from conftest import golden_lines

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def apg():
    import importlib

    m = importlib.import_module("apogee-vm_amd")
    m.default_context()
    return m


def _kat(orc):
    recs = list(golden_lines("mercury_proof.txt"))
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


def test_open_replays_committed_proof(apg, orc):
    k = _kat(orc)
    srs = apg.Srs.generate(k["tau"], 1 << k["nv"])
    cm = apg.g1_to_wire(apg.commit(srs, k["f"]))
    assert cm == k["cm"]
    tr = apg.Transcript()
    v, proof = apg.open(srs, k["f"], cm, k["u"], tr)
    assert orc.to_hex_le(orc.FR, v) == k["claim"]
    assert proof.hex() == k["proof"]
    assert orc.to_hex_le(orc.FR, tr.sample()) == k["probe"]


@pytest.fixture(scope="module")
def srs16(apg, orc):
    rng = orc.SplitMix(9090)
    tau = rng.next_fr()
    return tau, apg.Srs.generate(tau, 1 << 16), orc.srs_powers(tau, 1 << 16, workers=16)


@pytest.mark.parametrize("nv", [2, 4, 6, 8, 10, 12, 14, 16])
def test_open_matches_oracle(apg, orc, srs16, nv):
    _, srs, powers = srs16
    rng = orc.SplitMix(7000 + nv)
    f = np.stack([rng.next_fr() for _ in range(1 << nv)])
    u = np.stack([rng.next_fr() for _ in range(nv)])
    cm = apg.g1_to_wire(apg.commit(srs, f))
    tr_g, tr_o = apg.Transcript(), orc.Transcript()
    v, proof = apg.open(srs, f, cm, u, tr_g)
    v_o, proof_o = orc.mercury_open(powers, f, cm, u, tr_o, workers=16)
    assert np.array_equal(v, v_o)
    assert proof == proof_o
    assert np.array_equal(tr_g.sample(), tr_o.sample())


def test_open_device_table_and_witness_mix(apg, orc, srs16):
    """The device-table entry on a witness-like column (zeros, ones, small
    values, full-width values), against the host-table entry and the oracle."""
    _, srs, powers = srs16
    nv = 14
    ctx = apg.default_context()
    buf = apg.DeviceBuffer(ctx, 32 << nv)
    apg.synthetic_device(ctx, 4242, 1 << nv, "witness", out=buf)
    f = buf.download().reshape(-1, 4)
    rng = orc.SplitMix(55)
    u = np.stack([rng.next_fr() for _ in range(nv)])
    cm = apg.g1_to_wire(apg.commit(srs, f))
    v_d, p_d = apg.open(srs, buf, cm, u, apg.Transcript(), num_vars=nv)
    v_h, p_h = apg.open(srs, f, cm, u, apg.Transcript())
    v_o, p_o = orc.mercury_open(powers, f, cm, u, orc.Transcript(), workers=16)
    assert p_d == p_h == p_o
    assert np.array_equal(v_d, v_o) and np.array_equal(v_h, v_o)


def test_open_errors(apg, orc, srs16):
    _, srs, _ = srs16
    rng = orc.SplitMix(1)
    f = np.stack([rng.next_fr() for _ in range(1 << 5)])
    with pytest.raises(apg.UnsupportedNumVars):
        apg.open(srs, f, bytes(64), np.stack([rng.next_fr() for _ in range(5)]), apg.Transcript())
    f4 = f[:16]
    with pytest.raises(apg.PointLengthMismatch):
        apg.open(srs, f4, bytes(64), np.stack([rng.next_fr() for _ in range(3)]), apg.Transcript())
    small = apg.Srs.generate(rng.next_fr(), 8)
    with pytest.raises(apg.SrsTooSmall):
        apg.open(small, f4, bytes(64), np.stack([rng.next_fr() for _ in range(4)]), apg.Transcript())


def test_batch_open_replays_committed_proof(apg, orc):
    """batch_open of one column replays mercury_batch.txt: the preamble before
    the opening moves every challenge (crates/pcs/tests/vectors/mercury_batch.txt)."""
    recs = list(golden_lines("mercury_batch.txt"))
    one = lambda key: next(r for r in recs if r[0] == key)[1]
    nv = int(one("numvars"))
    vals = {int(r[1]): r[2] for r in recs if r[0] == "value"}
    pts = {int(r[1]): r[2] for r in recs if r[0] == "point"}
    f = np.stack([orc.from_hex_le(orc.FR, vals[i]) for i in range(1 << nv)])
    u = np.stack([orc.from_hex_le(orc.FR, pts[i]) for i in range(nv)])
    srs = apg.Srs.generate(orc.from_hex_le(orc.FR, one("tau")), 1 << nv)
    cm = bytes.fromhex(one("commitment"))
    tr = apg.Transcript()
    vs, proof = apg.batch_open(srs, [f], [cm], u, tr)
    assert orc.to_hex_le(orc.FR, vs[0]) == one("claim")
    assert proof.hex() == one("proof")
    assert orc.to_hex_le(orc.FR, tr.sample()) == one("probe")


@pytest.mark.parametrize("width,nr,shape", [(6, 0, [1, 1, 1]), (5, 1, [2, 1]), (4, 2, [4, 3, 1]),
                                             (11, 1, [2, 2, 1, 2])])
def test_batch_open_stacked_matches_oracle(apg, orc, srs16, width, nr, shape):
    _, srs, powers = srs16
    rng = orc.SplitMix(100 * width + nr)
    stacks = [[np.stack([rng.next_fr() for _ in range(1 << width)]) for _ in range(k)] for k in shape]
    cms = []
    for s in stacks:
        table = np.concatenate(s + [np.zeros((1 << width, 4), dtype=np.uint64)] * ((1 << nr) - len(s)))
        cms.append(apg.g1_to_wire(apg.commit(srs, table)))
    u = np.stack([rng.next_fr() for _ in range(width)])
    r = [rng.next_fr() for _ in range(nr)]
    tr_g, tr_o = apg.Transcript(), orc.Transcript()
    vs, proof = apg.batch_open_stacked(srs, stacks, cms, u, r, tr_g)
    vs_o, proof_o = orc.mercury_batch_open_stacked(powers, stacks, cms, u, r, tr_o, workers=16)
    assert np.array_equal(vs, vs_o)
    assert proof == proof_o
    assert np.array_equal(tr_g.sample(), tr_o.sample())


def test_batch_open_errors(apg, orc, srs16):
    _, srs, _ = srs16
    rng = orc.SplitMix(3)
    c = np.stack([rng.next_fr() for _ in range(16)])
    u = np.stack([rng.next_fr() for _ in range(4)])
    with pytest.raises(apg.EmptyBatch):
        apg.batch_open_stacked(srs, [], [], u, [], apg.Transcript())
    with pytest.raises(apg.BatchLengthMismatch):
        apg.batch_open(srs, [c, c], [bytes(64)], u, apg.Transcript())
    with pytest.raises(apg.MixedColumnSizes):
        apg.batch_open(srs, [c, c[:8]], [bytes(64)] * 2, u, apg.Transcript())
    c8 = c[:8]
    with pytest.raises(apg.BatchLengthMismatch):
        apg.batch_open_stacked(srs, [[c8, c8, c8]], [bytes(64)], u[:3], [rng.next_fr()], apg.Transcript())
    with pytest.raises(apg.UnsupportedNumVars):
        apg.batch_open_stacked(srs, [[c, c]], [bytes(64)], u, [rng.next_fr()], apg.Transcript())
