"""The oracle's GKR layer, pinned:
- the gate kernel against the known answers of crates/gkr/tests/kernel.rs
  (every shape with literal coefficients other than 0 and ±1, a challenge
  coefficient and nonzero constants, written out as integer arithmetic);
- each virtual table's closed form against its row values on the cube and
  its multilinearity in every coordinate (together they fix the extension);
- the flattened list (the package's ResolvedList) and the oracle's gate
  values against an integer interpreter of the shapes;
- the layer sumcheck's rounds, point and final values against the rounds'
  definition from multilinear extensions on small layers (row-wise with every
  shape, virtual table and a cached entry; halving with tree shapes);
- the forward pass against the interpreter row by row.
"""
import numpy as np
import pytest

import gkr_cases as gc
from gkr_cases import F, I, R, Fs, Is


@pytest.fixture(scope="module")
def g(orc):
    return gc.gkr_mod()


GAMMA = 0x9E3779B97F4A7C15 * 0x0123456789ABCDEF % R


def ev(orc, shape, coeffs, split, values):
    return I(orc.gkr_eval_gate(shape, Fs(coeffs), split, Fs(values)))


def test_eval_gate_kernel_kats(orc):
    """crates/gkr/tests/kernel.rs:35-189, shape by shape."""
    G = GAMMA
    # linear (kernel.rs:38-58): Σ c_i x_i + c_0, coefficients then constant
    assert ev(orc, orc.GATE_LINEAR, [2, 5, G, 7], 0, [11, 13, 17]) == (2 * 11 + 5 * 13 + 7 + G * 17) % R
    assert ev(orc, orc.GATE_LINEAR, [3, G], 0, [11]) == (33 + G) % R
    # product (kernel.rs:62-80)
    v = ev(orc, orc.GATE_PRODUCT, [3], 0, [11, 13])
    assert v == 3 * 11 * 13 and v != 11 * 13
    assert ev(orc, orc.GATE_PRODUCT, [G], 0, [11, 13]) == G * 143 % R
    # mask into identity (kernel.rs:84-97)
    assert ev(orc, orc.GATE_MASK, [], 0, [11, 1]) == 11
    assert ev(orc, orc.GATE_MASK, [], 0, [11, 0]) == 1
    assert ev(orc, orc.GATE_MASK, [], 0, [11, 5]) == (11 * 5 + 1 - 5) % R
    # affine product (kernel.rs:103-135): a_1..a_t, a_0, b_1..b_u, b_0
    v = ev(orc, orc.GATE_AFFINE_PRODUCT, [2, G, 3, 5, 7], 2, [11, 13, 17])
    left = (2 * 11 + 3 + G * 13) % R
    assert v == left * (5 * 17 + 7) % R
    assert v != left * (5 * 17 + 3) % R
    assert v != (2 * 11 + G * 13) * (5 * 17 + 7) % R
    assert ev(orc, orc.GATE_AFFINE_PRODUCT, [2, 3, 5, G], 1, [11, 17]) == (2 * 11 + 3) * (5 * 17 + G) % R
    # quadratic (kernel.rs:144-189): c_0, a_1..a_t, b_1..b_u; products y, z
    v = ev(orc, orc.GATE_QUADRATIC, [7, 2, G, 3, G, 5], 2, [11, 13, 17, 19, 23, 29, 11, 17])
    products = (3 * 17 * 19 + 5 * 11 * 17 + G * 23 * 29) % R
    assert v == (7 + 2 * 11 + G * 13 + products) % R
    assert v != (2 * 11 + G * 13 + products) % R
    assert v != (7 + 2 * 11 + 17 * 19 + 5 * 11 * 17 + G * 13 + G * 23 * 29) % R
    assert ev(orc, orc.GATE_QUADRATIC, [G, 3], 0, [11, 13]) == (G + 429) % R
    assert ev(orc, orc.GATE_QUADRATIC, [7, 2], 1, [11]) == 29
    assert ev(orc, orc.GATE_QUADRATIC, [7], 0, []) == 7
    # halving shapes (crates/gkr-verify/src/lib.rs:187-190): x0 x1; p0 q1 + p1 q0
    assert ev(orc, orc.GATE_TREE_PRODUCT, [], 0, [11, 13]) == 143
    assert ev(orc, orc.GATE_TREE_CROSS, [], 0, [2, 3, 5, 7]) == 2 * 7 + 3 * 5


def test_eval_gate_planted_mistake_is_seen(orc):
    """The kernel KATs see a dropped constant: the affine product's right
    factor with the left constant is another value."""
    wrong = ev(orc, orc.GATE_AFFINE_PRODUCT, [2, GAMMA, 3, 5, 3], 2, [11, 13, 17])
    right = ev(orc, orc.GATE_AFFINE_PRODUCT, [2, GAMMA, 3, 5, 7], 2, [11, 13, 17])
    assert wrong != right


@pytest.mark.parametrize("n", [1, 3, 8, 9, 14, 15, 17])
def test_virtual_tables_extend_their_rows(orc, n):
    """virtual_at_point at every sampled cube point equals virtual_at_row, and
    is affine in each coordinate at a random point: so it is the table's
    multilinear extension (docs/spec/gkr.md §2.1)."""
    rng = np.random.default_rng(n)
    rows = sorted(set([0, (1 << n) - 1] + [int(x) for x in rng.integers(0, 1 << n, 24)]))
    for kind in range(7):
        for y in rows:
            bits = Fs([(y >> j) & 1 for j in range(n)])
            assert I(orc.gkr_virtual_at_point(kind, bits)) == gc.virtual_row(kind, y), (kind, y)
            assert I(orc.gkr_virtual_at_row(kind, y)) == gc.virtual_row(kind, y)
        p = [int(x) for x in rng.integers(0, 2**62, n)]
        for j in range(n):
            def at(t):
                q = list(p)
                q[j] = t
                return I(orc.gkr_virtual_at_point(kind, Fs(q)))
            u = int(rng.integers(2, 2**62))
            f0, f1 = at(0), at(1)
            assert at(u) == (f0 + u * (f1 - f0)) % R, (kind, j)


def test_virtual_extension_small_cube_by_definition(orc):
    """At a random point of 5 variables every closed form equals
    Σ_y eq(z, y) row(y), the extension by definition."""
    rng = np.random.default_rng(7)
    z = [int(x) for x in rng.integers(0, 2**62, 5)]
    for kind in range(7):
        want = gc.mle([gc.virtual_row(kind, y) for y in range(32)], z)
        assert I(orc.gkr_virtual_at_point(kind, Fs(z))) == want


def _flat(orc, resolved):
    a = resolved.arrays()
    return orc.GkrFlat(a["gates"], a["ncached"], a["nproducing"], a["nenforcing"], a["sources"], a["coeffs"],
                       a["kinds"], a["halving"])


def _int_cols(rng, k, rows):
    return [[int(x) % R for x in rng.integers(0, 2**62, rows)] for _ in range(k)]


def test_resolved_gate_values_match_interpreter(orc, g):
    art, ncols = gc.rowwise_artifact()
    ch = gc.challenges()
    res = g.ResolvedList(art, 0, ch)
    flat = _flat(orc, res)
    rng = np.random.default_rng(11)
    n = 5
    cols = _int_cols(rng, ncols, 1 << n)
    for y in [0, 1, 7, 19, 31]:
        lower = Fs([c[y] for c in cols])
        virtuals = Fs([gc.virtual_row(int(kd), y) for kd, _ in art.virtuals])
        got = Is(orc.gkr_gate_values(flat, lower, None, virtuals))
        assert got == gc.row_values(art, 0, cols, y, n, ch)


def test_resolver_layout(g):
    """ResolvedList::new's order: cached, producing, enforcing; a halving
    gate reads each operand at both children (crates/gkr-verify/src/lib.rs:
    330-395)."""
    art, _ = gc.rowwise_artifact()
    res = g.ResolvedList(art, 0, gc.challenges())
    spec = art.layers[0]
    assert res.gates.shape[0] == len(spec.cached) + len(spec.producing) + len(spec.enforcing)
    assert (res.ncached, res.nproducing, res.nenforcing) == (1, len(spec.producing), 2)
    # p1 = Product(11, W0, W1): columns 1 and 2 of layout M, W, S
    p1 = res.gates[1 + 1]
    assert list(res.sources[p1[1]:p1[2]]) == [1, 2]
    art_h, _ = gc.halving_artifact()
    rh = g.ResolvedList(art_h, 1)
    t1 = rh.gates[1]  # TreeCross(L1, L0): L1 lower, L1 upper, L0 lower, L0 upper
    assert list(rh.sources[t1[1]:t1[2]]) == [1, (1 << 24) | 1, 0, 1 << 24]
    assert rh.kinds.shape[0] == 0
    with pytest.raises(g.GkrError):
        g.ResolvedList(art, 0, {})  # a challenge slot without a value


def test_forward_layer_rowwise(orc, g):
    art, ncols = gc.rowwise_artifact()
    ch = gc.challenges()
    flat = _flat(orc, g.ResolvedList(art, 0, ch))
    rng = np.random.default_rng(3)
    n = 6
    cols = _int_cols(rng, ncols, 1 << n)
    outs = orc.gkr_forward_layer(flat, [Fs(c) for c in cols], n)
    for y in range(1 << n):
        want = gc.row_values(art, 0, cols, y, n, ch)[:flat.nproducing]
        assert [I(o[y]) for o in outs] == want


def test_forward_layer_halving(orc, g):
    art, ncols = gc.halving_artifact()
    flat = _flat(orc, g.ResolvedList(art, 1))
    rng = np.random.default_rng(5)
    n = 4
    cols = _int_cols(rng, ncols, 1 << n)
    outs = orc.gkr_forward_layer(flat, [Fs(c) for c in cols], n)
    half = 1 << (n - 1)
    for y in range(half):
        ch0 = [c[y] for c in cols]
        ch1 = [c[y + half] for c in cols]
        want = [ch0[0] * ch1[0], ch0[1] * ch1[0] + ch1[1] * ch0[0], ch0[2] * ch1[3] + ch1[2] * ch0[3],
                ch0[3] * ch1[3]]
        assert [I(o[y]) for o in outs] == [w % R for w in want]


def _check_sumcheck(orc, g, art, k, lower, upper, n, seed):
    ch = gc.challenges()
    res = g.ResolvedList(art, k, ch)
    flat = _flat(orc, res)
    rng = np.random.default_rng(seed)
    r = [int(x) for x in rng.integers(0, 2**62, n)]
    w = [int(x) for x in rng.integers(1, 2**62, flat.nweighted)]
    t = orc.Transcript()
    rounds, point, finals = orc.gkr_prove_sumcheck(flat, Fs(w), Fs(r), [Fs(c) for c in lower],
                                                   [Fs(c) for c in upper] if upper else None, t)
    rho = Is(point)
    want = gc.brute_rounds(art, k, lower, upper or [], r, rho, w, ch)
    assert [Is(rd) for rd in rounds] == want
    assert Is(finals) == [gc.mle(c, rho) for c in lower + (upper or [])]
    # the verifier's identities (crates/gkr-verify): g_0(0) + g_0(1) = claim,
    # g_j(0) + g_j(1) = g_{j-1}(rho_{j-1})
    claim = sum(gc.eq_int(r, [(y >> b) & 1 for b in range(n)]) *
                gc.summand_at_point(art, k, lower, upper or [], [(y >> b) & 1 for b in range(n)], w, ch)
                for y in range(1 << n)) % R
    for j, c in enumerate(want):
        assert (2 * c[0] + c[1] + c[2] + c[3]) % R == claim
        x = rho[j]
        claim = (c[0] + x * (c[1] + x * (c[2] + x * c[3]))) % R


@pytest.mark.parametrize("n", [1, 2, 3])
def test_layer_sumcheck_rowwise_by_definition(orc, g, n):
    art, ncols = gc.rowwise_artifact()
    rng = np.random.default_rng(100 + n)
    cols = _int_cols(rng, ncols, 1 << n)
    _check_sumcheck(orc, g, art, 0, cols, None, n, 200 + n)


@pytest.mark.parametrize("n", [1, 3])
def test_layer_sumcheck_halving_by_definition(orc, g, n):
    art, ncols = gc.halving_artifact()
    rng = np.random.default_rng(300 + n)
    full = _int_cols(rng, ncols, 1 << (n + 1))
    half = 1 << n
    lower = [c[:half] for c in full]
    upper = [c[half:] for c in full]
    _check_sumcheck(orc, g, art, 1, lower, upper, n, 400 + n)


def test_wide_gates_gate_values_and_sumcheck(orc, g):
    """Wide gates, three cached entries read by several gates: gate values
    against the interpreter, and the sumcheck against the definition at n = 2."""
    art, ncols = gc.wide_artifact()
    ch = gc.challenges()
    flat = _flat(orc, g.ResolvedList(art, 0, ch))
    rng = np.random.default_rng(21)
    n = 2
    cols = _int_cols(rng, ncols, 1 << n)
    for y in range(1 << n):
        lower = Fs([c[y] for c in cols])
        virtuals = Fs([gc.virtual_row(int(kd), y) for kd, _ in art.virtuals])
        assert Is(orc.gkr_gate_values(flat, lower, None, virtuals)) == gc.row_values(art, 0, cols, y, n, ch)
    _check_sumcheck(orc, g, art, 0, cols, None, n, 77)
