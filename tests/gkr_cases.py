"""Shared cases for the GKR layer tests: sample artifacts (every gate shape,
cached entries, every virtual table, halving lists), an interpreter of the
gate shapes over Python integers written from the shape table of
docs/spec/gkr.md §3 (crates/constraints/src/lib.rs:310-360), the virtual
tables' row values (docs/spec/gkr.md §2.1), and a brute-force definition of
the layer sumcheck's round polynomials from multilinear extensions.

Test infrastructure: it imports the oracle only for Montgomery conversions.
"""
import importlib

import numpy as np

import oracle as orc

R = 21888242871839275222246405745257275088548364400416034343698204186575808495617
RAM_LIVE_BIT = 14

apg_pkg = None


def gkr_mod():
    return importlib.import_module("apogee-vm_amd.gkr")


def F(x: int) -> np.ndarray:
    """Python int -> Montgomery limbs."""
    return orc.from_bytes(orc.FR, (x % R).to_bytes(32, "little"))


def I(a) -> int:
    """Montgomery limbs -> Python int."""
    return int.from_bytes(orc.to_bytes(orc.FR, np.asarray(a, dtype=np.uint64).reshape(4)), "little")


def Fs(xs):
    return np.stack([F(x) for x in xs]) if len(xs) else np.zeros((0, 4), dtype=np.uint64)


def Is(a):
    return [I(x) for x in np.asarray(a, dtype=np.uint64).reshape(-1, 4)]


# ---- row values of the virtual tables (docs/spec/gkr.md §2.1)
def virtual_row(kind: int, y: int) -> int:
    if kind == 0:
        return y
    if kind == 1:
        return 1 if y >= 1 << RAM_LIVE_BIT else 0
    if kind == 2:
        return y % (1 << 19)
    if kind == 3:
        return y % (1 << 16)
    if kind == 4:
        return y & 0xFF
    if kind == 5:
        return (y >> 8) & 0xFF
    return (y & 0xFF) ^ ((y >> 8) & 0xFF)


# ---- multilinear extensions by definition
def eq_int(r, y):
    out = 1
    for a, b in zip(r, y):
        out = out * (a * b + (1 - a) * (1 - b)) % R
    return out


def eq_weights(z):
    """eq(z, w) for every w of the cube, w's bit j = variable j."""
    w = [1]
    for zj in z:
        w = [x * (1 - zj) % R for x in w] + [x * zj % R for x in w]
    return w


def mle(values, z):
    """Σ_w eq(z, w) values[w] over the cube of len(z) variables."""
    ws = eq_weights(z)
    return sum(a * b for a, b in zip(ws, values)) % R


# ---- an interpreter of the shapes over integers (per docs/spec/gkr.md §3)
def coef_int(c, challenges):
    g = gkr_mod()
    if isinstance(c, g.Literal):
        return I(np.array(c.value, dtype=np.uint64))
    return I(challenges[c.slot])


def py_gate(gate, val, challenges):
    """gate's value, val(address) giving each operand's integer; a halving
    shape's val returns the pair (child 0, child 1)."""
    g = gkr_mod()
    c = lambda k: coef_int(k, challenges)
    if isinstance(gate, g.Linear):
        return (c(gate.constant) + sum(c(k) * val(a) for k, a in gate.terms)) % R
    if isinstance(gate, g.Product):
        return c(gate.coeff) * val(gate.left) * val(gate.right) % R
    if isinstance(gate, g.MaskIntoIdentity):
        x, m = val(gate.input), val(gate.mask)
        return (x * m + 1 - m) % R
    if isinstance(gate, g.AffineProduct):
        left = c(gate.left_constant) + sum(c(k) * val(a) for k, a in gate.left)
        right = c(gate.right_constant) + sum(c(k) * val(a) for k, a in gate.right)
        return left * right % R
    if isinstance(gate, g.Quadratic):
        out = c(gate.constant) + sum(c(k) * val(a) for k, a in gate.linear)
        out += sum(c(b) * val(y) * val(z) for b, y, z in gate.products)
        return out % R
    if isinstance(gate, g.TreeProduct):
        x0, x1 = val(gate.input)
        return x0 * x1 % R
    if isinstance(gate, g.TreeCross):
        p0, p1 = val(gate.left)
        q0, q1 = val(gate.right)
        return (p0 * q1 + p1 * q0) % R
    raise TypeError(gate)


def lit(x):
    g = gkr_mod()
    return g.Literal.of(F(x))


# ---- sample artifacts
CH_SLOT = 3


def challenges():
    return {CH_SLOT: F(0x9E3779B97F4A7C15 * 0x0123456789ABCDEF)}


def rowwise_artifact(all_virtuals=True, cached=True, enforcing=True):
    """One row-wise list over M[0], W[0..2], S[0] and virtual tables, with
    every row-wise shape, literal and challenge coefficients, nonzero
    constants, and (optionally) a cached entry read by two gates."""
    g = gkr_mod()
    ch = g.Challenge(CH_SLOT)
    M0, W0, W1, W2, S0 = g.Memory(0), g.Witness(0), g.Witness(1), g.Witness(2), g.Setup(0)
    kinds = list(g.VirtualKind) if all_virtuals else [g.VirtualKind.RowIndex, g.VirtualKind.Xor8Out]
    virtuals = [(k, f"v{int(k)}") for k in kinds]
    V = lambda k: g.Virtual(k)
    cached_list = []
    C0 = None
    if cached:
        C0 = g.Cached(0, 0)
        cached_list = [("c0", C0, g.Linear([(lit(3), W0), (ch, S0)], lit(2)))]
    lin_src = C0 if cached else W2
    producing = [
        ("p0", g.Inner(1, 0), g.Linear([(lit(2), M0), (lit(5), W1), (ch, V(kinds[0]))], lit(7))),
        ("p1", g.Inner(1, 1), g.Product(lit(11), W0, W1)),
        ("p2", g.Inner(1, 2), g.MaskIntoIdentity(W2, M0)),
        ("p3", g.Inner(1, 3), g.AffineProduct([(lit(2), M0), (ch, lin_src)], lit(3), [(lit(5), W2), (lit(13), V(kinds[-1]))],
                                               lit(7))),
        ("p4", g.Inner(1, 4), g.Quadratic(lit(7), [(lit(2), W0), (ch, V(kinds[len(kinds) // 2]))],
                                          [(lit(3), W1, W2), (ch, S0, lin_src), (lit(5), M0, M0)])),
        ("p5", g.Inner(1, 5), g.Product(ch, lin_src, V(kinds[1]))),
    ]
    for q, k in enumerate(kinds):
        producing.append((f"pv{q}", g.Inner(1, 6 + q), g.Product(lit(q + 2), V(k), W0)))
    enf = []
    if enforcing:
        enf = [("e0", g.Quadratic(lit(1), [(lit(4), S0)], [(lit(9), W0, W1)])),
               ("e1", g.AffineProduct([], lit(5), [(ch, W2)], lit(0)))]
    layer = g.LayerSpec(False, 0, len(producing), cached=cached_list, producing=producing, enforcing=enf)
    art = g.CircuitArtifact(["m0"], ["w0", "w1", "w2"], ["s0"], virtuals, [layer])
    return art, 5  # five committed columns


def halving_artifact():
    """A halving list over four columns: tree products, tree crosses (a
    numerator reading its denominator), and a column read twice."""
    g = gkr_mod()
    L = lambda j: g.Inner(1, j)
    producing = [
        ("t0", g.Inner(2, 0), g.TreeProduct(L(0))),
        ("t1", g.Inner(2, 1), g.TreeCross(L(1), L(0))),
        ("t2", g.Inner(2, 2), g.TreeCross(L(2), L(3))),
        ("t3", g.Inner(2, 3), g.TreeProduct(L(3))),
    ]
    layer0 = g.LayerSpec(False, 0, 4, producing=[("x", L(j), g.Linear([(lit(1), g.Witness(j))], lit(0))) for j in range(4)])
    layer1 = g.LayerSpec(True, 0, 4, producing=producing)
    art = g.CircuitArtifact([], ["w0", "w1", "w2", "w3"], [], [], [layer0, layer1])
    return art, 4


def row_values(art, k, cols, y, n, challenges_):
    """Every weighted gate of row-wise list k at row y (cached entries
    substituted), from the integer columns."""
    g = gkr_mod()
    spec = art.layers[k]
    m, w = len(art.memory), len(art.witness)
    cache = {}

    def val(a):
        if isinstance(a, g.Memory):
            return cols[a.i][y]
        if isinstance(a, g.Witness):
            return cols[m + a.i][y]
        if isinstance(a, g.Setup):
            return cols[m + w + a.i][y]
        if isinstance(a, g.Inner):
            return cols[a.offset][y]
        if isinstance(a, g.Virtual):
            return virtual_row(int(a.kind), y)
        return cache[a.offset]

    for q, (_, _, gate) in enumerate(spec.cached):
        cache[q] = py_gate(gate, val, challenges_)
    gates = [e[-1] for e in spec.producing] + [e[-1] for e in spec.enforcing]
    return [py_gate(gate, val, challenges_) for gate in gates]


def summand_at_point(art, k, tables_lo, tables_up, z, weights, challenges_):
    """S_k at the point z: each table's extension at z, virtual tables'
    extensions at z, cached entries evaluated there, gates weighted."""
    g = gkr_mod()
    spec = art.layers[k]
    m, w = len(art.memory), len(art.witness)
    n = len(z)
    lo = [mle(t, z) for t in tables_lo]
    up = [mle(t, z) for t in tables_up]
    vir = {int(kind): mle([virtual_row(int(kind), y) for y in range(1 << n)], z) for kind, _ in art.virtuals}
    cache = {}

    def col(a):
        if isinstance(a, g.Memory):
            return a.i
        if isinstance(a, g.Witness):
            return m + a.i
        if isinstance(a, g.Setup):
            return m + w + a.i
        return a.offset

    def val(a):
        if isinstance(a, g.Virtual):
            return vir[int(a.kind)]
        if isinstance(a, g.Cached):
            return cache[a.offset]
        if spec.halving:
            return (lo[col(a)], up[col(a)])
        return lo[col(a)]

    for q, (_, _, gate) in enumerate(spec.cached):
        cache[q] = py_gate(gate, val, challenges_)
    gates = [e[-1] for e in spec.producing] + [e[-1] for e in spec.enforcing]
    return sum(wt * py_gate(gate, val, challenges_) for wt, gate in zip(weights, gates)) % R


def brute_rounds(art, k, tables_lo, tables_up, r, rho, weights, challenges_):
    """g_j(X) at X = 0..3 from the definition, with the bound challenges rho:
    g_j(X) = Σ_y eq(r, (rho_<j, X, y)) S((rho_<j, X, y)), then the cubic's
    ascending coefficients by Newton's forward differences."""
    n = len(r)
    out = []
    inv2, inv6 = pow(2, R - 2, R), pow(6, R - 2, R)
    for j in range(n):
        vals = []
        for X in range(4):
            acc = 0
            for y in range(1 << (n - j - 1)):
                z = list(rho[:j]) + [X] + [(y >> b) & 1 for b in range(n - j - 1)]
                acc += eq_int(r, z) * summand_at_point(art, k, tables_lo, tables_up, z, weights, challenges_)
            vals.append(acc % R)
        d1 = (vals[1] - vals[0]) % R
        d2 = (vals[2] - 2 * vals[1] + vals[0]) % R
        d3 = (vals[3] - 3 * vals[2] + 3 * vals[1] - vals[0]) % R
        # v0 + d1 X + d2 X(X-1)/2 + d3 X(X-1)(X-2)/6, expanded
        c3 = d3 * inv6 % R
        c2 = (d2 * inv2 - 3 * c3) % R
        c1 = (d1 - d2 * inv2 + 2 * c3) % R
        out.append([vals[0], c1, c2, c3])
    return out


def wide_artifact(ncols=24, nterms=40):
    """A row-wise list with wide gates: a Linear of nterms terms, an
    AffineProduct of 10 + 10 terms, a Quadratic of 8 linear terms and 12
    products, three cached entries (each read by several gates), every
    column read, and two virtual tables."""
    g = gkr_mod()
    ch = g.Challenge(CH_SLOT)
    W = [g.Witness(i) for i in range(ncols)]
    V0, V1 = g.Virtual(g.VirtualKind.Range16), g.Virtual(g.VirtualKind.Xor8Out)
    cached = [
        ("c0", g.Cached(0, 0), g.Linear([(lit(3 + q), W[q]) for q in range(6)], lit(11))),
        ("c1", g.Cached(0, 1), g.Linear([(ch, W[7]), (lit(2), V0)], lit(0))),
        ("c2", g.Cached(0, 2), g.Quadratic(lit(5), [(lit(1), W[8])], [(lit(2), W[9], W[10])])),
    ]
    C0, C1, C2 = g.Cached(0, 0), g.Cached(0, 1), g.Cached(0, 2)
    producing = [
        ("lin", g.Inner(1, 0), g.Linear([(lit(q + 1) if q % 3 else ch, W[q % ncols]) for q in range(nterms)]
                                        + [(lit(7), C0), (lit(9), C2)], lit(13))),
        ("aff", g.Inner(1, 1), g.AffineProduct([(lit(q + 2), W[(q + 3) % ncols]) for q in range(9)] + [(ch, C1)],
                                               lit(4), [(lit(q + 5), W[(q + 11) % ncols]) for q in range(10)],
                                               lit(6))),
        ("quad", g.Inner(1, 2), g.Quadratic(lit(2), [(lit(q + 1), W[(q + 12) % ncols]) for q in range(7)] + [(ch, C2)],
                                            [(lit(q + 3), W[(2 * q) % ncols], W[(2 * q + 5) % ncols])
                                             for q in range(11)] + [(ch, C0, C1)])),
        ("mask", g.Inner(1, 3), g.MaskIntoIdentity(C0, W[ncols - 1])),
        ("prod", g.Inner(1, 4), g.Product(ch, V1, C1)),
    ]
    enf = [("e0", g.Linear([(lit(1), C0), (lit(2), C1), (lit(3), C2), (ch, V1)], lit(1)))]
    layer = g.LayerSpec(False, 0, len(producing), cached=cached, producing=producing, enforcing=enf)
    art = g.CircuitArtifact([], [f"w{i}" for i in range(ncols)], [], [(g.VirtualKind.Range16, "r16"),
                                                                     (g.VirtualKind.Xor8Out, "x8")], [layer])
    return art, ncols
