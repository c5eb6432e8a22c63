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

    # commitments of the first and last column against the oracle MSM
    bases = srs22.g1(0, n)
    for c in (0, k - 1):
        assert p.commitments[c] == orc.jac_to_wire(orc.msm(bases, cols[c], threads=8, workers=16))

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

    # round chain: g_j(0) + g_j(1) = g_{j-1}(rho_{j-1})
    one = orc.from_u64(orc.FR, 1)
    zero = orc.fe()
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
    col_at = [orc.evaluate(cols[c], chal) for c in range(k)]
    w = orc.from_u64(orc.FR, 1)
    for o in range(f):
        acc = orc.fe()
        for c in range(k):
            acc = orc.add(orc.FR, acc, orc.mul(orc.FR, w, col_at[c]))
            w = orc.mul(orc.FR, w, rho_w)
        assert np.array_equal(acc, p.final_evals[o]), f"table {o}"
