"""GPU parity for the sumcheck rounds and the shard composition, through the C
ABI, against the oracle (bit-exact round polynomials, points, final
evaluations and transcript state)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def apg():
    import importlib

    m = importlib.import_module("apogee-vm_amd")
    m.default_context()
    return m


def _gates(apg, orc):
    one = orc.from_u64(orc.FR, 1)
    m1 = orc.sub(orc.FR, orc.fe(), one)
    seven = orc.from_u64(orc.FR, 7)
    spec = {
        "product": (2, [(one, [0, 1])]),
        "product3": (3, [(one, [0, 1, 2])]),
        "square_minus": (2, [(one, [0, 0]), (m1, [1])]),
        "wide": (5, [(one, [0, 1]), (one, [0, 2]), (m1, [3, 4])]),
        "linear": (1, [(seven, [0])]),
        "mixed3": (4, [(seven, [0, 1, 2]), (m1, [3]), (one, [1, 3])]),
    }
    out = {}
    for name, (k, terms) in spec.items():
        g = apg.Gate(list(range(k)), [apg.GateTerm(c, *(fs + [None] * (3 - len(fs)))) for c, fs in terms])
        out[name] = (k, terms, g)
    return out


def _tables(orc, k, n, seed, witness=False):
    return [(orc.fill_witness if witness else orc.fill_uniform)(seed + t, 1 << n) for t in range(k)]


@pytest.mark.parametrize("gate", ["product", "product3", "square_minus", "wide", "linear", "mixed3"])
@pytest.mark.parametrize("n", [1, 2, 3, 5, 8])
def test_prove_sumcheck_vs_oracle(apg, orc, gate, n):
    k, terms, g = _gates(apg, orc)[gate]
    tabs = _tables(orc, k, n, 1000 * n + len(gate))
    eqp = orc.fill_uniform(77 + n, n)
    t_gpu, t_orc = apg.Transcript(), orc.Transcript()
    r1, p1, f1 = apg.prove_sumcheck(eqp, g, tabs, t_gpu)
    r2, p2, f2 = orc.prove_sumcheck(eqp, terms, tabs, t_orc)
    assert np.array_equal(r1, r2)
    assert np.array_equal(p1, p2)
    assert np.array_equal(f1, np.stack([x[0] for x in f2]))
    assert np.array_equal(t_gpu.challenge_scalar(5), t_orc.challenge(5))


@pytest.mark.parametrize("n,k,seed", [(12, 2, 5), (14, 3, 6)])
def test_prove_sumcheck_product_large(apg, orc, n, k, seed):
    one = orc.from_u64(orc.FR, 1)
    terms = [(one, list(range(k)))]
    g = apg.Gate.product(k)
    tabs = _tables(orc, k, n, seed, witness=True)
    eqp = orc.fill_uniform(seed, n)
    t_gpu, t_orc = apg.Transcript(), orc.Transcript()
    r1, p1, f1 = apg.prove_sumcheck(eqp, g, tabs, t_gpu)
    r2, p2, f2 = orc.prove_sumcheck(eqp, terms, tabs, t_orc, workers=8)
    assert np.array_equal(r1, r2)
    assert np.array_equal(p1, p2)
    assert np.array_equal(f1, np.stack([x[0] for x in f2]))


def test_zerocheck_vs_oracle(apg, orc):
    one = orc.from_u64(orc.FR, 1)
    m1 = orc.sub(orc.FR, orc.fe(), one)
    n = 7
    a = orc.fill_uniform(3, 1 << n)
    b = np.stack([orc.mul(orc.FR, v, v) for v in a])  # satisfies a*a - b = 0
    terms = [(one, [0, 0]), (m1, [1])]
    g = apg.Gate([0, 1], [apg.GateTerm(one, 0, 0), apg.GateTerm(m1, 1)])
    t_gpu, t_orc = apg.Transcript(), orc.Transcript()
    digest = orc.witness_digest([a, b])
    t_gpu.append_scalar(orc.TAG_WITNESS_DIGEST, digest)
    t_orc.append_scalars(orc.TAG_WITNESS_DIGEST, digest)
    proof = apg.prove_zerocheck(g, [a, b], t_gpu)
    r2, f2 = orc.prove_zerocheck(terms, [a, b], t_orc)
    assert np.array_equal(proof.rounds, r2)
    assert np.array_equal(proof.final_evals, f2)
    assert t_gpu.state()[:96] == t_orc.snapshot()[:96]
    # round 0 sums to zero on a satisfying witness
    s = orc.add(orc.FR, r2[0][0], orc.add(orc.FR, orc.add(orc.FR, r2[0][0], r2[0][1]), orc.add(orc.FR, r2[0][2], r2[0][3])))
    assert not s.any()


@pytest.mark.parametrize("nv,k,degree", [(10, 4, 3), (9, 3, 4), (12, 8, 3)])
def test_prove_shard_vs_oracle(apg, orc, nv, k, degree):
    n = 1 << nv
    tau = orc.fill_uniform(555, 1)[0]
    srs = apg.Srs.generate(tau, n)
    srs_aff = srs.g1()
    cols = np.stack([orc.fill_witness(7000 + c, n) for c in range(k)])
    ctx = apg.default_context()
    dev = apg.DeviceBuffer.from_array(ctx, cols)
    t_gpu, t_orc = apg.Transcript(), orc.Transcript()
    proof = apg.prove_shard(srs, dev, nv, k, degree, t_gpu)
    cms, rounds, fin = orc.shard_script(srs_aff, cols, degree, t_orc)
    assert proof.commitments == cms
    assert np.array_equal(proof.rounds, rounds)
    assert np.array_equal(proof.final_evals, fin)
    assert t_gpu.state()[:96] == t_orc.snapshot()[:96]


def test_config1_end_to_end(apg, orc):
    """One 2^16 column committed, then a 16-variable degree-3 sumcheck on
    eq(r, x) a(x) b(x) with r seeded: commitment + 16 rounds bit-exact."""
    n, nv = 1 << 16, 16
    seed = 2026
    tau = orc.fill_uniform(seed, 1)[0]
    srs = apg.Srs.generate(tau, n)
    srs_aff = srs.g1()
    f, a, b = (orc.fill_uniform(seed + 1 + i, n) for i in range(3))
    r = orc.fill_uniform(seed + 4, nv)
    one = orc.from_u64(orc.FR, 1)
    # oracle
    t_orc = orc.Transcript()
    cm_o = orc.jac_to_wire(orc.msm(srs_aff, f, workers=8))
    t_orc.append_g1(orc.TAG_COMMITMENT, [cm_o])
    r_o, p_o, f_o = orc.prove_sumcheck(r, [(one, [0, 1])], [a, b], t_orc, workers=8)
    # gpu
    t_gpu = apg.Transcript()
    cm_g = apg.g1_to_wire(apg.commit(srs, f))
    t_gpu.append_g1_points(orc.TAG_COMMITMENT, [cm_g])
    r_g, p_g, f_g = apg.prove_sumcheck(r, apg.Gate.product(2), [a, b], t_gpu)
    assert cm_g == cm_o
    assert np.array_equal(r_g, r_o)
    assert np.array_equal(p_g, p_o)
    assert np.array_equal(f_g, np.stack([x[0] for x in f_o]))


@pytest.mark.parametrize("k,nv", [(1, 0), (3, 6), (5, 12)])
def test_witness_digest_dev_matches_oracle(apg, orc, k, nv):
    """apg_witness_digest_dev over device-resident columns (k x 2^nv,
    column-major) equals the oracle's digest and the host entry's."""
    ctx = apg.default_context()
    cols = [orc.fill_witness(600 + c, 1 << nv) for c in range(k)]
    buf = apg.DeviceBuffer.from_array(ctx, np.concatenate(cols))
    out = np.zeros(4, dtype=np.uint64)
    assert apg.lib().apg_witness_digest_dev(ctx.h, buf.ptr, k, nv, out.ctypes.data) == 0
    assert np.array_equal(out, orc.witness_digest(cols))
    assert np.array_equal(out, apg.witness_digest(cols))
    buf.free()


def _lift(orc, col):
    if col.dtype == np.bool_:
        col = col.astype(np.uint32)
    return np.stack([orc.from_u64(orc.FR, int(v)) for v in col])


@pytest.mark.parametrize("n,kinds,shape", [
    (1, ("u1", "u8", "fr"), "product3"),   # every round on the host, from the caller's tables
    (5, ("u8", "u16"), "product2"),        # likewise
    (7, ("u1", "u16", "u32"), "product3"),  # device round 0 through the lift, the host folds T^0 next
    (13, ("u8", "u32", "fr"), "product3"),  # device rounds 0 (lift) and 1 (fold through the lift)
    (12, ("u16", "u1"), "gate"),           # the general gate kernel through the lift
])
def test_lifted_backings_match_oracle(apg, orc, n, kinds, shape):
    """Narrow tables (PolyBacking U1/U8/U16/U32, crates/poly/src/lib.rs:46-97)
    kept at their width on the device and lifted in the first rounds give the
    oracle's proof over the lifted Fr tables."""
    ctx = apg.default_context()
    rng = np.random.default_rng(700 + n)
    cols = []
    for i, kd in enumerate(kinds):
        m = 1 << n
        if kd == "u1":
            cols.append(rng.integers(0, 2, m).astype(np.bool_))
        elif kd == "u8":
            cols.append(rng.integers(0, 256, m, dtype=np.uint8))
        elif kd == "u16":
            cols.append(rng.integers(0, 1 << 16, m, dtype=np.uint16))
        elif kd == "u32":
            cols.append(rng.integers(0, 1 << 32, m, dtype=np.uint32))
        else:
            cols.append(orc.fill_uniform(710 + i, m))
    lifted = [c if c.ndim == 2 else _lift(orc, c) for c in cols]
    one = orc.from_u64(orc.FR, 1)
    if shape == "gate":
        gate = apg.Gate([0, 1], [apg.GateTerm(apg.fr_from_u64(3), 0, 1), apg.GateTerm(one, 1)])
        oterms = [(orc.from_u64(orc.FR, 3), [0, 1]), (one, [1])]
    else:
        k = len(kinds)
        gate = apg.Gate.product(k)
        oterms = [(one, list(range(k)))]
    r = orc.fill_uniform(720 + n, n)
    dev = [apg.upload_backed(ctx, c) for c in cols]
    assert [kd for _, kd in dev] == list(kinds)
    rg, pg, fg = apg.prove_sumcheck_device(r, gate, dev, apg.Transcript())
    ro, po, fo = orc.prove_sumcheck(r, oterms, lifted, orc.Transcript())
    assert np.array_equal(rg, ro) and np.array_equal(pg, po)
    assert np.array_equal(fg, np.stack([x[0] for x in fo]) if isinstance(fo, list) else fo.reshape(fg.shape))
    for b, _ in dev:
        b.free()


def _column(orc, rng, kind, m, seed):
    if kind == "u8":
        return rng.integers(0, 256, m, dtype=np.uint8)
    if kind == "u16":
        return rng.integers(0, 1 << 16, m, dtype=np.uint16)
    if kind == "u32":
        return rng.integers(0, 1 << 32, m, dtype=np.uint32)
    return orc.fill_uniform(seed, m)


# Two-table products run their device rounds two per launch after the single
# ones (round 0, the lifted rounds, and one more when the count of device
# rounds left is odd); the host takes over once the tables hold at most 256
# entries in all (round n - 7 here).
@pytest.mark.parametrize("n,zero_at,kinds", [
    (9, None, ("fr", "fr")),    # device rounds 0-1: no pair
    (10, None, ("fr", "fr")),   # round 0, then rounds 1-2 paired
    (11, None, ("fr", "fr")),   # rounds 0 and 1 alone, then 2-3
    (13, None, ("fr", "fr")),   # rounds 0 and 1 alone, then 2-3 and 4-5
    (12, 3, ("fr", "fr")),      # r_3 = 0 on a device round: every round alone, round 3 on all nodes
    (12, 8, ("fr", "fr")),      # r_8 = 0 on a host round: the device rounds still paired
    (12, None, ("u8", "u32")),  # lifted rounds 0-1 and round 2 alone, then 3-4
    (13, None, ("u16", "fr")),  # lifted rounds 0-1 alone, then 2-3 and 4-5
])
def test_paired_rounds_match_oracle(apg, orc, n, zero_at, kinds):
    ctx = apg.default_context()
    rng = np.random.default_rng(900 + n)
    cols = [_column(orc, rng, kd, 1 << n, 910 + 10 * n + i) for i, kd in enumerate(kinds)]
    lifted = [c if c.ndim == 2 else _lift(orc, c) for c in cols]
    r = orc.fill_uniform(930 + n, n)
    if zero_at is not None:
        r[zero_at] = 0
    one = orc.from_u64(orc.FR, 1)
    dev = [apg.upload_backed(ctx, c) for c in cols]
    t_gpu, t_orc = apg.Transcript(), orc.Transcript()
    rg, pg, fg = apg.prove_sumcheck_device(r, apg.Gate.product(2), dev, t_gpu)
    ro, po, fo = orc.prove_sumcheck(r, [(one, [0, 1])], lifted, t_orc, workers=8)
    assert np.array_equal(rg, ro) and np.array_equal(pg, po)
    assert np.array_equal(fg, np.stack([x[0] for x in fo]))
    assert np.array_equal(t_gpu.challenge_scalar(5), t_orc.challenge(5))
    for b, _ in dev:
        b.free()
