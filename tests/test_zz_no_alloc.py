"""apg_sumcheck_prove_dev_noalloc through the C ABI: the same rounds, point,
final evaluations and transcript state as apg_sumcheck_prove_dev on the same
device tables, on calls in a row that shrink (the device's scratch array
keeps the previous call's contents), and a clear refusal above its limit."""
import ctypes as C
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def apg():
    m = importlib.import_module("apogee-vm_amd")
    m.default_context()
    return m


def _prove(apg, name, gate, bufs, n, r):
    L = apg.lib()
    ctx = apg.default_context()
    ptrs = (C.c_void_p * len(bufs))(*[b.ptr.value for b in bufs])
    d = gate.degree()
    rounds = np.zeros((n, d + 1, 4), dtype=np.uint64)
    point = np.zeros((n, 4), dtype=np.uint64)
    finals = np.zeros((len(bufs), 4), dtype=np.uint64)
    t = apg.Transcript()
    cb = C.cast(L.apg_transcript_round, apg.ROUND_FN)
    rc = getattr(L, name)(ctx.h, gate.c_terms(), len(gate.terms), ptrs, len(bufs), n,
                          np.ascontiguousarray(r).ctypes.data, cb, t.h, rounds.ctypes.data, point.ctypes.data,
                          finals.ctypes.data)
    return rc, rounds, point, finals, t.state()


def _tables(apg, k, n, seed):
    ctx = apg.default_context()
    return [apg.synthetic_device(ctx, seed + t, 1 << n, "witness") for t in range(k)]


# (gate, tables, log2 rows): largest first, each call smaller than the last;
# the first is at the limit (1 x 2^21 entries); product2 at 2^12 runs its
# device rounds two per launch, at 2^5 all on the host
CASES = [("product1", 1, 21), ("product3", 3, 18), ("product2", 2, 12), ("two_terms", 2, 11), ("product2", 2, 5)]


def _gate(apg, name):
    one = apg.fr_from_u64(1)
    if name.startswith("product"):
        return apg.Gate.product(int(name[-1]))
    return apg.Gate([0, 1], [apg.GateTerm(apg.fr_from_u64(5), 0, 1), apg.GateTerm(one, 1)])


def test_noalloc_equals_the_allocating_entry(apg):
    for i, (gname, k, n) in enumerate(CASES):
        gate = _gate(apg, gname)
        r = apg.synthetic(7000 + i, n, "uniform")
        a = _tables(apg, k, n, 100 * i)
        b = _tables(apg, k, n, 100 * i)
        want = _prove(apg, "apg_sumcheck_prove_dev", gate, a, n, r)
        got = _prove(apg, "apg_sumcheck_prove_dev_noalloc", gate, b, n, r)
        print(f"case {gname} k={k} n={n}: rc {want[0]} / {got[0]}", flush=True)
        assert want[0] == 0 and got[0] == 0, apg.lib().apg_last_error()
        for x, y in zip(want[1:4], got[1:4]):
            assert np.array_equal(x, y), gname
        assert want[4] == got[4]
        for buf in a + b:
            buf.free()


def test_noalloc_refuses_above_its_limit(apg):
    n, k = 21, 2  # 2^22 entries: 96 MiB of fold buffers
    gate = apg.Gate.product(2)
    bufs = _tables(apg, k, n, 900)
    r = apg.synthetic(901, n, "uniform")
    rc = _prove(apg, "apg_sumcheck_prove_dev_noalloc", gate, bufs, n, r)[0]
    msg = apg.lib().apg_last_error().decode()
    assert rc == -2 and "limit" in msg, msg
    for buf in bufs:
        buf.free()
