"""GPU parity for the general GKR layer through the C ABI: the device's layer
sumcheck (rounds, point, bound tables, transcript state) and forward pass
against the oracle, bit-exact, on row-wise lists with every shape, every
virtual table and a cached entry, and on halving lists; plus the refusals."""
import numpy as np
import pytest

import gkr_cases as gc
from gkr_cases import Fs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def apg():
    import importlib

    m = importlib.import_module("apogee-vm_amd")
    m.default_context()
    return m


@pytest.fixture(scope="module")
def g(apg):
    return gc.gkr_mod()


def _flat(orc, resolved):
    a = resolved.arrays()
    return orc.GkrFlat(a["gates"], a["ncached"], a["nproducing"], a["nenforcing"], a["sources"], a["coeffs"],
                       a["kinds"], a["halving"])


def _tables(orc, k, n, seed):
    return [orc.fill_ctr(seed + t, 1 << n, "uniform") for t in range(k)]


def _run_both(apg, g, orc, art, k, lower, upper, n, seed):
    ch = gc.challenges()
    res = g.ResolvedList(art, k, ch)
    flat = _flat(orc, res)
    rng = np.random.default_rng(seed)
    r = Fs([int(x) for x in rng.integers(0, 2**62, n)])
    w = Fs([int(x) for x in rng.integers(1, 2**62, flat.nweighted)])
    t_o = orc.Transcript()
    t_o.append_scalars(orc.TAG_SUMCHECK_CHALLENGE, r[:1] if n else orc.fe(1))
    rounds_o, point_o, finals_o = orc.gkr_prove_sumcheck(flat, w, r, lower, upper, t_o)
    t_d = apg.Transcript()
    t_d.append_scalars(orc.TAG_SUMCHECK_CHALLENGE, r[:1] if n else orc.fe(1))
    rounds_d, point_d, finals_d = g.prove_sumcheck(r, res, w, lower, upper, t_d)
    print(f"gkr sumcheck n={n} halving={res.halving} tables={len(lower) + len(upper or [])}", flush=True)
    assert np.array_equal(rounds_d, rounds_o)
    assert np.array_equal(point_d, point_o)
    assert np.array_equal(finals_d, finals_o)
    assert t_d.state()[:96] == t_o.snapshot()[:96]


@pytest.mark.parametrize("n", [1, 2, 5, 9, 12])
def test_layer_sumcheck_rowwise(apg, g, orc, n):
    art, ncols = gc.rowwise_artifact()
    _run_both(apg, g, orc, art, 0, _tables(orc, ncols, n, 10 * n), None, n, n)


def test_layer_sumcheck_rowwise_ram_live_and_range(apg, g, orc):
    """n = 16: RamLive's high variables and Range16/Xor8 read real bits."""
    art, ncols = gc.rowwise_artifact()
    _run_both(apg, g, orc, art, 0, _tables(orc, ncols, 16, 77), None, 16, 16)


def test_layer_sumcheck_rowwise_plain(apg, g, orc):
    """No cached entry, no enforcing gates, two virtual tables."""
    art, ncols = gc.rowwise_artifact(all_virtuals=False, cached=False, enforcing=False)
    _run_both(apg, g, orc, art, 0, _tables(orc, ncols, 10, 5), None, 10, 3)


@pytest.mark.parametrize("n", [1, 4, 11])
def test_layer_sumcheck_halving(apg, g, orc, n):
    art, ncols = gc.halving_artifact()
    full = _tables(orc, ncols, n + 1, 50 + n)
    half = 1 << n
    lower = [np.ascontiguousarray(c[:half]) for c in full]
    upper = [np.ascontiguousarray(c[half:]) for c in full]
    _run_both(apg, g, orc, art, 1, lower, upper, n, 60 + n)


def test_layer_sumcheck_no_variables(apg, g, orc):
    art, ncols = gc.rowwise_artifact()
    res = g.ResolvedList(art, 0, gc.challenges())
    cols = _tables(orc, ncols, 0, 9)
    w = Fs([1] * (res.nproducing + res.nenforcing))
    rounds, point, finals = g.prove_sumcheck([], res, w, cols, None, apg.Transcript())
    assert rounds.shape[0] == 0 and point.shape[0] == 0
    assert np.array_equal(finals, np.stack([c[0] for c in cols]))


def test_layer_sumcheck_device_tables(apg, g, orc):
    """The _dev entry on HBM-resident tables gives the host entry's proof."""
    art, ncols = gc.rowwise_artifact()
    n = 8
    cols = _tables(orc, ncols, n, 123)
    res = g.ResolvedList(art, 0, gc.challenges())
    rng = np.random.default_rng(8)
    r = Fs([int(x) for x in rng.integers(0, 2**62, n)])
    w = Fs([int(x) for x in rng.integers(1, 2**62, res.nproducing + res.nenforcing)])
    want = g.prove_sumcheck(r, res, w, cols, None, apg.Transcript())
    ctx = apg.default_context()
    bufs = [apg.DeviceBuffer.from_array(ctx, c) for c in cols]
    got = g.prove_sumcheck_device(r, res, w, bufs, None, apg.Transcript())
    for a, b in zip(got, want):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("n", [1, 7, 15])
def test_forward_rowwise(apg, g, orc, n):
    art, ncols = gc.rowwise_artifact()
    res = g.ResolvedList(art, 0, gc.challenges())
    cols = _tables(orc, ncols, n, 500 + n)
    want = orc.gkr_forward_layer(_flat(orc, res), cols, n)
    got = g.forward_layer(res, cols, n)
    print(f"gkr forward n={n}", flush=True)
    for a, b in zip(got, want):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("n", [1, 10])
def test_forward_halving_and_whole_circuit(apg, g, orc, n):
    art, ncols = gc.halving_artifact()
    cols = _tables(orc, ncols, n, 700 + n)
    layers = g.forward(art, cols)
    l1 = orc.gkr_forward_layer(_flat(orc, g.ResolvedList(art, 0)), cols, n)
    l2 = orc.gkr_forward_layer(_flat(orc, g.ResolvedList(art, 1)), l1, n)
    assert len(layers) == 2
    for a, b in zip(layers[0] + layers[1], l1 + l2):
        assert np.array_equal(a, b)


def test_forward_device_buffers(apg, g, orc):
    art, ncols = gc.rowwise_artifact()
    res = g.ResolvedList(art, 0, gc.challenges())
    n = 9
    cols = _tables(orc, ncols, n, 31)
    want = g.forward_layer(res, cols, n)
    ctx = apg.default_context()
    ins = [apg.DeviceBuffer.from_array(ctx, c) for c in cols]
    outs = [apg.DeviceBuffer(ctx, (1 << n) * 32) for _ in range(res.nproducing)]
    g.forward_layer_device(res, ins, n, outs)
    for b, w in zip(outs, want):
        assert np.array_equal(b.download().reshape(-1, 4), w)


def test_refusals(apg, g, orc):
    """Malformed lists are APG_ERR_ARG before any kernel runs."""
    art, ncols = gc.rowwise_artifact()
    res = g.ResolvedList(art, 0, gc.challenges())
    cols = _tables(orc, ncols, 3, 1)
    w = Fs([1] * (res.nproducing + res.nenforcing))
    r = Fs([5, 6, 7])
    a = res.arrays()
    # a degree-3 gate: Product of a cached degree-2 entry and a column
    gates = a["gates"].copy()
    gates[0] = (1, gates[0][1], gates[0][1] + 2, 0, 0)  # cached entry as Product(c, col, col)
    bad = g.ResolvedList.from_arrays(gates, a["ncached"], a["nproducing"], a["nenforcing"], a["sources"],
                                     a["coeffs"], a["kinds"], False)
    with pytest.raises(apg.ApgError):
        g.prove_sumcheck(r, bad, w, cols, None, apg.Transcript())
    # a column source past the columns
    src = a["sources"].copy()
    src[0] = 99
    bad = g.ResolvedList.from_arrays(a["gates"], a["ncached"], a["nproducing"], a["nenforcing"], src, a["coeffs"],
                                     a["kinds"], False)
    with pytest.raises(apg.ApgError):
        g.forward_layer(bad, cols, 3)
    # a child-1 source in a row-wise list
    src = a["sources"].copy()
    src[0] = (1 << 24) | 0
    bad = g.ResolvedList.from_arrays(a["gates"], a["ncached"], a["nproducing"], a["nenforcing"], src, a["coeffs"],
                                     a["kinds"], False)
    with pytest.raises(apg.ApgError):
        g.forward_layer(bad, cols, 3)
    # coefficients past the array
    bad = g.ResolvedList.from_arrays(a["gates"], a["ncached"], a["nproducing"], a["nenforcing"], a["sources"],
                                     a["coeffs"][:2], a["kinds"], False)
    with pytest.raises(apg.ApgError):
        g.forward_layer(bad, cols, 3)


def test_full_size_layer_properties(apg, g, orc):
    """2^20 rows (the bench layer's size) with the every-shape list: the
    forward pass equal to the oracle's, and the sumcheck's size-independent
    identities: g_j(0) + g_j(1) = g_(j-1)(rho_(j-1)) for every round, the
    transcript replayed from the rounds gives the point, and the last round
    at its challenge equals eq(r, rho) * S(rho), S evaluated by the oracle on
    the bound tables and the virtual tables' closed forms at rho."""
    art, ncols = gc.rowwise_artifact()
    ch = gc.challenges()
    res = g.ResolvedList(art, 0, ch)
    flat = _flat(orc, res)
    n = 20
    cols = [orc.fill_ctr(900 + t, 1 << n, "witness") for t in range(ncols)]
    fwd = g.forward_layer(res, cols, n)
    want = orc.gkr_forward_layer(flat, cols, n)
    for a, b in zip(fwd, want):
        assert np.array_equal(a, b)
    rng = np.random.default_rng(20)
    r = Fs([int(x) for x in rng.integers(0, 2**62, n)])
    w = Fs([int(x) for x in rng.integers(1, 2**62, flat.nweighted)])
    t = apg.Transcript()
    rounds, point, finals = g.prove_sumcheck(r, res, w, cols, None, t)
    print("gkr full size n=20", flush=True)
    R = gc.R
    rs, rho = [gc.Is(x) for x in rounds], gc.Is(point)
    replay = orc.Transcript()
    claim = None
    for j, c in enumerate(rs):
        if claim is not None:
            assert (2 * c[0] + c[1] + c[2] + c[3]) % R == claim
        replay.append_scalars(orc.TAG_SUMCHECK_ROUND, rounds[j])
        assert gc.I(replay.challenge(orc.TAG_SUMCHECK_CHALLENGE)) == rho[j]
        x = rho[j]
        claim = (c[0] + x * (c[1] + x * (c[2] + x * c[3]))) % R
    virt = np.stack([orc.gkr_virtual_at_point(int(k), point) for k in flat.kinds])
    s = gc.I(orc.gkr_summand(flat, w, finals, None, virt))
    e = gc.eq_int(gc.Is(r), rho)
    assert claim == e * s % R


@pytest.mark.parametrize("n", [3, 10, 14])
def test_wide_gates_and_cached_entries(apg, g, orc, n):
    """Gates of 40 operands, three cached entries, two virtual tables: the
    sumcheck (its small rounds split over (pair, gate) items) and the forward
    pass against the oracle."""
    art, ncols = gc.wide_artifact()
    cols = _tables(orc, ncols, n, 3000 + n)
    _run_both(apg, g, orc, art, 0, cols, None, n, 40 + n)
    res = g.ResolvedList(art, 0, gc.challenges())
    want = orc.gkr_forward_layer(_flat(orc, res), cols, n)
    for a, b in zip(g.forward_layer(res, cols, n), want):
        assert np.array_equal(a, b)
