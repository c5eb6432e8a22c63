"""The lazy 29-bit Fq arithmetic and the bucket accumulation chain as compiled
for gfx950, run in a kernel and checked against the oracle (the host build of
the same code is checked in test_native_host.py)."""
import numpy as np
import pytest

from test_native_host import accumulate29_cases, accumulate29_run, fq29_cases, fq29_run

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def apg():
    import importlib

    return importlib.import_module("apogee-vm_amd")


@pytest.fixture(scope="module")
def ctx(apg):
    c = apg.Context(0)
    yield c
    c.close()


def test_fq29_on_device_matches_oracle(apg, orc, ctx):
    a, b, want = fq29_cases(orc)
    for op in range(6):
        got = fq29_run(apg, ctx.h, op, a, b)
        bad = np.nonzero((got != want[op]).any(axis=1))[0]
        assert bad.size == 0, (op, bad[:8].tolist())


@pytest.mark.parametrize("form", [0, 1, 2])
def test_accumulate_chains_on_device_match_oracle(apg, orc, ctx, form):
    """form 0 is the lazy 29-bit bucket loop the MSM runs, form 1 the 32-bit
    one, form 2 the lazy general XYZZ add of the bucket reduction."""
    points, neg, offsets, want = accumulate29_cases(orc)
    got = accumulate29_run(apg, ctx.h, form, points, neg, offsets)
    bad = [c for c in range(want.shape[0]) if not np.array_equal(got[c], want[c])]
    assert not bad, bad[:8]
