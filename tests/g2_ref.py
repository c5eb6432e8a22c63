"""BN254's G2 over Python integers, for the SRS tests: Fq2 = Fq[u]/(u^2 + 1),
the twist y^2 = x^3 + 3/(9 + u), affine addition and doubling, a square
root in Fq2 (q = 3 mod 4), and the canonical 128-byte wire form
x.c0 || x.c1 || y.c0 || y.c1 (docs/spec/srs.md §4). The generator is the one
EIP-197 publishes. Test infrastructure only."""

Q = 21888242871839275222246405745257275088696311157297823662689037894645226208583
R = 21888242871839275222246405745257275088548364400416034343698204186575808495617

GEN = (
    (10857046999023057135944570762232829481370756359578518086990519993285655852781,
     11559732032986387107991004021392285783925812861821192530917403151452391805634),
    (8495653923123431417604973247489272438418190587263600148770280649306958101930,
     4082367875863433681332203403145435568316851327593401208105741076214120093531),
)


def add(a, b):
    return ((a[0] + b[0]) % Q, (a[1] + b[1]) % Q)


def sub(a, b):
    return ((a[0] - b[0]) % Q, (a[1] - b[1]) % Q)


def mul(a, b):
    return ((a[0] * b[0] - a[1] * b[1]) % Q, (a[0] * b[1] + a[1] * b[0]) % Q)


def inv(a):
    n = pow(a[0] * a[0] + a[1] * a[1], Q - 2, Q)
    return (a[0] * n % Q, -a[1] * n % Q)


def pw(a, e):
    out = (1, 0)
    while e:
        if e & 1:
            out = mul(out, a)
        a = mul(a, a)
        e >>= 1
    return out


B = mul((3, 0), inv((9, 1)))


def on_curve(p):
    x, y = p
    return mul(y, y) == add(mul(mul(x, x), x), B)


def double(p):
    x, y = p
    lam = mul(mul((3, 0), mul(x, x)), inv(add(y, y)))
    x3 = sub(mul(lam, lam), add(x, x))
    return (x3, sub(mul(lam, sub(x, x3)), y))


def add_pts(p, q):
    if p is None:
        return q
    if q is None:
        return p
    if p[0] == q[0]:
        return double(p) if p[1] == q[1] else None
    lam = mul(sub(q[1], p[1]), inv(sub(q[0], p[0])))
    x3 = sub(sub(mul(lam, lam), p[0]), q[0])
    return (x3, sub(mul(lam, sub(p[0], x3)), p[1]))


def scalar(p, k):
    acc = None
    while k:
        if k & 1:
            acc = add_pts(acc, p)
        p = double(p) if p is not None else None
        k >>= 1
    return acc


def sqrt(a):
    """A square root in Fq2 for q = 3 mod 4 (the complex method), or None."""
    a1 = pw(a, (Q - 3) // 4)
    alpha = mul(mul(a1, a1), a)
    a0 = mul(pw(alpha, Q), alpha)
    if a0 == (Q - 1, 0):
        return None
    x0 = mul(a1, a)
    if alpha == (Q - 1, 0):
        return mul((0, 1), x0)
    b = pw(add((1, 0), alpha), (Q - 1) // 2)
    return mul(b, x0)


def point_off_subgroup(seed: int):
    """A point of the twist outside the order-r subgroup: x from the seed, y a
    root of x^3 + b'; the cofactor is ~2^254, so r P = O does not hold."""
    x = (seed % Q, (seed * 7 + 3) % Q)
    while True:
        rhs = add(mul(mul(x, x), x), B)
        y = sqrt(rhs)
        if y is not None and mul(y, y) == rhs:
            p = (x, y)
            if scalar(p, R) is not None:
                return p
        x = ((x[0] + 1) % Q, x[1])


def wire(p) -> bytes:
    if p is None:
        return bytes(128)
    (x0, x1), (y0, y1) = p
    return b"".join(v.to_bytes(32, "little") for v in (x0, x1, y0, y1))
