"""The sumcheck captured as one HIP graph with the transcript on the device
(apg_sumcheck_prove_graph_dev): the same rounds, point, final evaluations and
transcript state as the host-driven device path (itself checked against the
oracle), for 1-3 tables from 1 to 2^16 entries and seeded transcripts, on
calls in a row; one case against the oracle directly."""
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def apg():
    m = importlib.import_module("apogee-vm_amd")
    m.default_context()
    return m


def _case(apg, orc, k, n, seed):
    ctx = apg.default_context()
    tabs = [orc.fill_uniform(seed * 10 + t, 1 << n) for t in range(k)]
    bufs = [apg.DeviceBuffer.from_array(ctx, a) for a in tabs]
    r = orc.fill_uniform(seed * 10 + 9, n)
    t = apg.Transcript()
    t.append_bytes(2, b"graph %d" % seed)
    return tabs, bufs, r, t


@pytest.mark.parametrize("k,n", [(1, 1), (2, 1), (1, 7), (2, 5), (3, 3), (2, 12), (3, 10), (2, 16)])
def test_graph_sumcheck_matches_host_driven_path(apg, orc, k, n):
    for seed in (n * 7 + k, n * 7 + k + 1):  # two calls in a row, different inputs
        tabs, bufs, r, t = _case(apg, orc, k, n, seed)
        t2 = t.clone()
        ra, pa, fa = apg.prove_sumcheck_device(r, apg.Gate.product(k), [(b, "fr") for b in bufs], t)
        rb, pb, fb = apg.prove_sumcheck_graph(r, bufs, t2)
        assert np.array_equal(ra, rb) and np.array_equal(pa, pb) and np.array_equal(fa, fb), (k, n, seed)
        assert t.state() == t2.state()
        # the tables are read, not modified
        for a, b in zip(tabs, bufs):
            assert np.array_equal(b.download().reshape(-1, 4), a)


def test_graph_sumcheck_matches_oracle(apg, orc):
    k, n = 2, 9
    tabs, bufs, r, t = _case(apg, orc, k, n, 4242)
    seed_state = t.state()
    rb, pb, fb = apg.prove_sumcheck_graph(r, bufs, t)
    ot = orc.Transcript()
    ot.append_bytes(2, b"graph 4242")
    one = orc.from_u64(orc.FR, 1)
    ro, po, fo = orc.prove_sumcheck(r, [(one, [0, 1])], tabs, ot)
    assert np.array_equal(rb, ro) and np.array_equal(pb, po)
    assert np.array_equal(fb, np.stack([np.asarray(f).reshape(4) for f in fo]))
    assert t.state()[:96] == ot.snapshot()[:96] and seed_state != t.state()


def test_graph_sumcheck_refusals(apg, orc):
    import ctypes as C

    L = apg.lib()
    ctx = apg.default_context()
    _, bufs, r, t = _case(apg, orc, 2, 4, 1)
    ptrs = (C.c_void_p * 4)(*([b.ptr.value for b in bufs] * 2))
    out = np.zeros((64, 4), dtype=np.uint64)
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    rr = np.ascontiguousarray(r)
    assert L.apg_sumcheck_prove_graph_dev(ctx.h, ptrs, 4, 4, p(rr), t.h, p(out), p(out), p(out)) == -2  # k = 4
    assert L.apg_sumcheck_prove_graph_dev(ctx.h, ptrs, 0, 4, p(rr), t.h, p(out), p(out), p(out)) == -2
    assert L.apg_sumcheck_prove_graph_dev(ctx.h, ptrs, 2, 0, p(rr), t.h, p(out), p(out), p(out)) == -2
