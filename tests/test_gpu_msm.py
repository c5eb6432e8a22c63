"""GPU parity for the MSM and commit path, through the C ABI, against the
reference's MSM KATs and the oracle (bit-exact affine points)."""
import numpy as np
import pytest

from conftest import golden_lines
from test_oracle_curve import msm_case

pytestmark = pytest.mark.gpu

P_FR = 0x30644E72E131A029B85045B68181585D2833E84879B9709143E1F593F0000001


@pytest.fixture(scope="module")
def apg():
    import importlib

    m = importlib.import_module("apogee-vm_amd")
    m.default_context()
    return m


def wire(apg, mont: bytes) -> str:
    return apg.g1_to_wire(mont).hex()


def test_msm_kats(apg, orc):
    for rec in golden_lines("msm_kats.txt"):
        _, pattern, n, seed, want = rec
        bases, scalars = msm_case(orc, pattern, int(n), int(seed))
        assert wire(apg, apg.msm(bases, scalars)) == want, (pattern, n)


def _edge_scalars(orc, n, seed):
    rng = orc.SplitMix(seed)
    s = np.stack([rng.next_fr() for _ in range(n)])
    pm1 = orc.from_bytes(orc.FR, (P_FR - 1).to_bytes(32, "little"))
    for i in range(0, n, 7):
        s[i] = 0
    for i in range(1, n, 11):
        s[i] = orc.from_u64(orc.FR, 1)
    for i in range(2, n, 13):
        s[i] = pm1
    return s


@pytest.mark.parametrize("n", [1, 2, 3, 31, 32, 33, 100, 1000, 4097])
def test_msm_vs_oracle(apg, orc, n):
    bases, _ = msm_case(orc, "random", n, 900 + n)
    if n > 4:
        bases[3] = 0  # infinity
        bases[4] = bases[5]  # repeated point (P + P in one bucket)
        bases[6] = bases[5]
        bases[6, 4:] = orc.neg(orc.FQ, bases[5, 4:])  # P and -P
    s = _edge_scalars(orc, n, 77 + n)
    if n > 6:
        s[5] = s[4]
        s[6] = s[4]
    want = orc.jac_to_wire(orc.msm(bases, s, workers=8)).hex()
    assert wire(apg, apg.msm(bases, s)) == want


@pytest.mark.parametrize("n", [1, 33, 1024, 5000])
def test_msm_small_u32_vs_oracle(apg, orc, n):
    bases, _ = msm_case(orc, "random", n, 1234 + n)
    rng = np.random.default_rng(n)
    s = rng.integers(0, 2**32, size=n, dtype=np.uint64).astype(np.uint32)
    s[::5] = 0xFFFFFFFF
    s[1::9] = 0
    s[2::17] = np.uint32(1) << (np.arange(len(s[2::17])) % 32).astype(np.uint32)
    want = orc.jac_to_wire(orc.msm_small_u32(bases, s, workers=8)).hex()
    assert wire(apg, apg.msm_small_u32(bases, s)) == want


def test_heavy_buckets(apg, orc):
    """All-one and all-(p-1) scalars pile every digit into one bucket."""
    n = 1 << 14
    bases, _ = msm_case(orc, "random", n, 31337)
    ones = np.repeat(orc.from_u64(orc.FR, 1)[None], n, axis=0)
    want = orc.jac_to_wire(orc.msm(bases, ones, workers=8)).hex()
    assert wire(apg, apg.msm(bases, ones)) == want
    pm1 = orc.from_bytes(orc.FR, (P_FR - 1).to_bytes(32, "little"))
    neg = np.repeat(pm1[None], n, axis=0)
    w2 = orc.jac_to_wire(orc.msm(bases, neg, workers=8))
    assert wire(apg, apg.msm(bases, neg)) == w2.hex()
    # -(sum) and (sum) differ only in y
    assert w2[:32].hex() == want[:64]


@pytest.fixture(scope="module")
def srs12(apg, orc):
    rng = orc.SplitMix(4242)
    tau = rng.next_fr()
    s = apg.Srs.generate(tau, 1 << 12)
    return tau, s


def test_srs_generate_matches_oracle(apg, orc, srs12):
    tau, s = srs12
    want = orc.srs_powers(tau, 1 << 12, workers=8)
    got = s.g1()
    assert np.array_equal(got, want)


def test_mercury_commitment_kat(apg, orc):
    recs = list(golden_lines("mercury_proof.txt"))
    tau = orc.from_hex_le(orc.FR, next(r for r in recs if r[0] == "tau")[1])
    nv = int(next(r for r in recs if r[0] == "numvars")[1])
    vals = {int(r[1]): r[2] for r in recs if r[0] == "value"}
    f = np.stack([orc.from_hex_le(orc.FR, vals[i]) for i in range(1 << nv)])
    srs = apg.Srs.generate(tau, 1 << nv)
    cm = apg.commit(srs, f)
    assert wire(apg, cm) == next(r for r in recs if r[0] == "commitment")[1]


@pytest.mark.parametrize("n", [1 << 12, 1 << 11, 3000, 1 << 6])
def test_commit_batch_witness_mix(apg, orc, srs12, n):
    _, s = srs12
    srs_aff = s.g1()
    cols = [orc.fill_witness(500 + n + c, n) for c in range(5)]
    got = apg.commit_batch(s, cols)
    for c in range(5):
        want = orc.jac_to_wire(orc.msm(srs_aff[:n], cols[c], workers=8)).hex()
        assert wire(apg, got[c]) == want, (n, c)


def test_commit_narrow_backings(apg, orc, srs12):
    _, s = srs12
    srs_aff = s.g1()
    n = 1 << 10
    rng = np.random.default_rng(7)
    for dtype in (np.uint8, np.uint16, np.uint32, np.bool_):
        col = rng.integers(0, 2**32, size=n, dtype=np.uint64).astype(dtype)
        want = orc.jac_to_wire(orc.msm_small_u32(srs_aff[:n], col.astype(np.uint32), workers=8)).hex()
        assert wire(apg, apg.commit(s, col)) == want, dtype
        lifted = np.stack([orc.from_u64(orc.FR, int(v)) for v in col.astype(np.uint64)])
        assert wire(apg, apg.commit(s, lifted)) == want


def test_commit_2_16_uniform(apg, orc):
    """config-1 size: one 2^16 uniform column over a 2^16 SRS."""
    n = 1 << 16
    tau = orc.fill_uniform(99, 1)[0]
    s = apg.Srs.generate(tau, n)
    srs_aff = s.g1()
    # spot-check the device SRS on the CPU
    rng = np.random.default_rng(1)
    for i in list(rng.integers(0, n, size=8)) + [0, 1, n - 1]:
        p = orc.fe()
        p[:] = orc.from_u64(orc.FR, 1)
        e = np.array([int(i), 0, 0, 0], dtype=np.uint64)
        ti = orc.pow_(orc.FR, tau, e)
        want = orc.gen_mul_batch(ti[None], workers=1)[0]
        assert np.array_equal(srs_aff[int(i)], want), i
    f = orc.fill_uniform(100, n)
    want = orc.jac_to_wire(orc.msm(srs_aff, f, workers=8)).hex()
    assert wire(apg, apg.commit(s, f)) == want


@pytest.mark.parametrize("n", [1 << 12, 3000])
def test_msm_tuning_knobs_keep_commitments(apg, orc, srs12, n):
    """The autotuner's knobs (window width, slices per resident accumulation
    thread) change the schedule, never the points: 8 witness-mix columns
    committed under several settings of both equal the oracle's."""
    _, s = srs12
    srs_aff = s.g1()
    cols = [orc.fill_witness(4000 + n + i, n) for i in range(8)]
    want = [orc.jac_to_wire(orc.msm(srs_aff[:n], c, workers=8)).hex() for c in cols]
    lg = n.bit_length() - 1
    try:
        for c, rounds in ((0, 1), (0, 3), (0, 8), (0, 16), (8, 2), (11, 6)):
            apg.set_window_bits(lg, c)
            apg.set_slice_rounds(lg, rounds)
            assert apg.slice_rounds(n) == rounds
            got = apg.commit_batch(s, cols)
            assert [wire(apg, g) for g in got] == want, (c, rounds)
    finally:
        apg.set_window_bits(lg, 0)
        apg.set_slice_rounds(lg, 0)
    assert apg.slice_rounds(n) == 4


def test_msm_tuning_knob_refusals(apg):
    L = apg.lib()
    assert L.apg_msm_set_slice_rounds(10, 17) == -2
    assert L.apg_msm_set_slice_rounds(10, -1) == -2
    assert L.apg_msm_set_slice_rounds(64, 4) == -2
    assert L.apg_msm_set_window_bits(10, 18) == -2
