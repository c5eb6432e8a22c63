"""The benchmark's full-size single shards on the GPU, checked at full size.

* 2^20 rows x 8 witness-like columns, degree 3 (the single large shard):
  every byte against the oracle's restatement of the whole shard script.
* 2^22 rows x 16 witness-like columns, degree 4 (the largest shard): through
  properties that do not need the oracle to redo the whole shard: two column
  commitments against the oracle MSM; the transcript replayed from the proof
  gives the same challenges; every round satisfies g_j(0) + g_j(1) =
  g_{j-1}(rho_{j-1}) (crates/sumcheck/src/lib.rs:441); the last round's value
  at its challenge equals eq(r, rho) * prod_o t_o(rho); and the final
  evaluations equal the combined columns evaluated at rho by the oracle.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def apg():
    import importlib

    m = importlib.import_module("apogee-vm_amd")
    m.default_context()
    return m


@pytest.fixture(scope="module")
def srs22(apg):
    tau = apg.synthetic(20261203, 1, "uniform")[0]
    srs = apg.Srs.generate(tau, 1 << 22, apg.default_context())
    yield srs
    srs.release()


def _columns(apg, nv, k, seed):
    ctx = apg.default_context()
    buf = apg.DeviceBuffer(ctx, 32 * k << nv)
    for c in range(k):
        apg.synthetic_device(ctx, seed + c, 1 << nv, "witness", out=buf, offset_elems=c << nv)
    return buf


def _horner(orc, coeffs, x):
    acc = coeffs[-1].copy()
    for c in coeffs[-2::-1]:
        acc = orc.add(orc.FR, orc.mul(orc.FR, acc, x), c)
    return acc


def test_single_large_shard_matches_oracle(apg, orc, srs22):
    nv, k, degree = 20, 8, 3
    buf = _columns(apg, nv, k, 4242)
    t = apg.Transcript()
    p = apg.prove_shard(srs22, buf, nv, k, degree, t)
    cols = buf.download().reshape(k, 1 << nv, 4)
    ot = orc.Transcript()
    cms, rounds, fin = orc.shard_script(srs22.g1(0, 1 << nv), cols, degree, ot, workers=16)
    assert p.commitments == cms
    assert np.array_equal(p.rounds, rounds) and np.array_equal(p.final_evals, fin)
    assert t.state()[:96] == ot.snapshot()[:96]  # sponge state words
    buf.free()


def test_largest_shard_properties(apg, orc, srs22):
    nv, k, degree = 22, 16, 4
    n, f = 1 << nv, degree - 1
    buf = _columns(apg, nv, k, 777)
    p = apg.prove_shard(srs22, buf, nv, k, degree, apg.Transcript())
    assert p.rounds.shape == (nv, degree + 1, 4) and p.final_evals.shape == (f, 4)
    cols = buf.download().reshape(k, n, 4)
    buf.free()

    # commitments against the oracle MSM: column 0 directly, and every column
    # through a random linear combination (sum_c g^c cm_c = MSM(bases,
    # sum_c g^c col_c) by linearity; the combination's own check fails for a
    # wrong column with probability 1 - 16/r)
    bases = srs22.g1(0, n)
    assert p.commitments[0] == orc.jac_to_wire(orc.msm(bases, cols[0], threads=8, workers=16))
    gam = orc.fill_uniform(4242, 1)[0]
    comb, lhs, w = orc.fe(n), None, orc.from_u64(orc.FR, 1)
    for c in range(k):
        orc.axpy(w, cols[c], comb)
        a = orc.affine_from_wire(p.commitments[c])
        assert a is not None
        term = orc.jac_mul(orc.jac_from_affine(a), w)
        lhs = term if lhs is None else orc.jac_add(lhs, term)
        w = orc.mul(orc.FR, w, gam)
    assert orc.jac_to_wire(lhs) == orc.jac_to_wire(orc.msm(bases, comb, threads=8, workers=16))
    del comb

    # the transcript replayed from the proof's own messages
    t = orc.Transcript()
    t.append_g1(orc.TAG_COMMITMENT, p.commitments)
    rho_w = t.challenge(orc.TAG_MERCURY_BATCH)
    r = np.stack([t.challenge(orc.TAG_SUMCHECK_CHALLENGE) for _ in range(nv)])
    chal = []
    for j in range(nv):
        t.append_scalars(orc.TAG_SUMCHECK_ROUND, p.rounds[j])
        chal.append(t.challenge(orc.TAG_SUMCHECK_CHALLENGE))
    chal = np.stack(chal)

    # the combined tables t_o = sum_c rho_w^(o k + c) col_c, and round 0
    # against the claim: g_0(0) + g_0(1) = sum_x eq(r, x) t_0(x) t_1(x) t_2(x)
    pw = [orc.from_u64(orc.FR, 1)]
    for _ in range(f * k - 1):
        pw.append(orc.mul(orc.FR, pw[-1], rho_w))
    tabs = [orc.fe(n) for _ in range(f)]
    for c in range(k):
        for o in range(f):
            orc.axpy(pw[o * k + c], cols[c], tabs[o])
    one = orc.from_u64(orc.FR, 1)
    zero = orc.fe()
    g00 = orc.add(orc.FR, _horner(orc, p.rounds[0], zero), _horner(orc, p.rounds[0], one))
    assert np.array_equal(g00, orc.product_claim(r, tabs)), "round 0 against the claim"

    # round chain: g_j(0) + g_j(1) = g_{j-1}(rho_{j-1})
    for j in range(1, nv):
        g0 = _horner(orc, p.rounds[j], zero)
        g1 = _horner(orc, p.rounds[j], one)
        prev = _horner(orc, p.rounds[j - 1], chal[j - 1])
        assert np.array_equal(orc.add(orc.FR, g0, g1), prev), f"round {j}"

    # last round at its challenge = eq(r, rho) * prod of the final evaluations
    last = _horner(orc, p.rounds[nv - 1], chal[nv - 1])
    prod = orc.eq_eval(r, chal)
    for o in range(f):
        prod = orc.mul(orc.FR, prod, p.final_evals[o])
    assert np.array_equal(last, prod)

    # final evaluations = the rho_w-combined columns evaluated at rho
    for o in range(f):
        assert np.array_equal(orc.evaluate(tabs[o], chal), p.final_evals[o]), f"table {o}"
