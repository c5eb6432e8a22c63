"""ctypes binding of the CPU oracle (``oracle/liboracle.so``).

TEST INFRASTRUCTURE ONLY: tests/, the smoke entry and bench.py's
cpu_baseline leg import this module as the checker. The product package never
imports it.

Field elements cross as numpy ``uint64`` arrays of shape ``(..., 4)`` holding
the reference's in-memory Montgomery limbs; affine G1 points as ``(..., 8)``
(Montgomery x || y, all-zero = infinity); Jacobian points as ``(..., 12)``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

FR, FQ = 0, 1

TAG_COMMITMENT = 3
TAG_SUMCHECK_ROUND = 4
TAG_SUMCHECK_CHALLENGE = 5
TAG_EVALUATION_CLAIM = 6
TAG_WITNESS_DIGEST = 8
TAG_SUMCHECK_FINAL_EVALS = 9
TAG_MERCURY_BATCH = 17


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


P = C.c_void_p
SZ = C.c_size_t
U64 = C.c_uint64


def _declare(L):
    def d(name, args, res=None):
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res

    for n in ("orc_mul", "orc_add", "orc_sub"):
        d(n, [C.c_int, P, P, P])
    d("orc_neg", [C.c_int, P, P])
    d("orc_square", [C.c_int, P, P])
    d("orc_inverse", [C.c_int, P, P], C.c_int)
    d("orc_pow", [C.c_int, P, P, P])
    d("orc_from_u64", [C.c_int, U64, P])
    d("orc_to_bytes", [C.c_int, P, P])
    d("orc_from_bytes", [C.c_int, P, P], C.c_int)
    d("orc_g1_add", [P, P, P])
    d("orc_g1_add_affine", [P, P, P])
    d("orc_g1_double", [P, P])
    d("orc_g1_mul", [P, P, P])
    d("orc_g1_mul_le_bytes", [P, P, P])
    d("orc_g1_to_affine", [P, P])
    d("orc_g1_from_affine", [P, P])
    d("orc_g1_generator", [P])
    d("orc_g1_affine_to_wire", [P, P])
    d("orc_g1_affine_from_wire", [P, P], C.c_int)
    d("orc_msm_window_bits", [SZ], SZ)
    d("orc_msm", [P, P, SZ, SZ, SZ, P], C.c_int)
    d("orc_msm_small_u32", [P, P, SZ, SZ, SZ, P], C.c_int)
    d("orc_srs_powers", [P, SZ, SZ, P])
    d("orc_g1_gen_mul_batch", [P, SZ, SZ, P])
    d("orc_bind", [P, SZ, P])
    d("orc_eq_table", [P, SZ, P])
    d("orc_eq_eval", [P, P, SZ, P])
    d("orc_evaluate", [P, SZ, P, P])
    d("orc_poseidon2_permute", [P])
    d("orc_tr_init", [P])
    d("orc_tr_observe", [P, P])
    d("orc_tr_sample", [P, P])
    d("orc_tr_append_scalars", [P, U64, P, SZ])
    d("orc_tr_append_bytes", [P, U64, P, SZ])
    d("orc_tr_challenge", [P, U64, P])
    d("orc_g1_limbs", [P, P])
    d("orc_tr_append_g1", [P, U64, P, SZ])
    d("orc_interpolate", [P, SZ, P])
    d("orc_prove_zerocheck", [P, SZ, P, SZ, SZ, P, P, P])
    d("orc_mercury_open", [P, SZ, P, SZ, P, P, P, SZ, P, P], C.c_int)
    d("orc_mercury_batch_open_stacked", [P, SZ, P, P, SZ, SZ, P, P, P, SZ, P, SZ, P, P], C.c_int)
    d("orc_prove_sumcheck", [P, SZ, P, SZ, P, SZ, P, SZ, P, P])
    d("orc_witness_digest", [P, SZ, SZ, P])
    d("orc_combine", [P, SZ, SZ, P, SZ, P])
    d("orc_axpy", [SZ, P, P, P])
    d("orc_msm_recode", [P, SZ, P], SZ)
    d("orc_msm_nonzero_digits", [P, SZ, SZ], C.c_uint64)
    d("orc_product_claim", [P, SZ, P, SZ, P])
    d("orc_splitmix_next", [P], U64)
    d("orc_splitmix_next_fr", [P, P])
    d("orc_chacha_block", [P, C.c_uint32, P, P])
    d("orc_fill_uniform", [U64, P, SZ])
    d("orc_fill_witness", [U64, P, SZ])
    d("orc_fill_ctr", [U64, C.c_int, P, SZ])
    U32 = C.c_uint32
    d("orc_gkr_eval_gate", [U32, P, U32, P, U32, P])
    d("orc_gkr_virtual_at_row", [U32, U64, P])
    d("orc_gkr_virtual_at_point", [U32, P, SZ, P])
    d("orc_gkr_gate_values", [P, P, SZ, P, SZ, P, P])
    d("orc_gkr_summand", [P, P, P, SZ, P, SZ, P, P])
    d("orc_gkr_forward_layer", [P, P, SZ, SZ, P])
    d("orc_gkr_prove_sumcheck", [P, P, P, SZ, P, P, SZ, P, P, P])


def ptr(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


def fe(n=None):
    return np.zeros((4,) if n is None else (n, 4), dtype=np.uint64)


# ---------------------------------------------------------------- fields
def from_bytes(which: int, b: bytes):
    out = fe()
    buf = np.frombuffer(bytes(b), dtype=np.uint8).copy()
    ok = lib().orc_from_bytes(which, ptr(buf), ptr(out))
    return out if ok else None


def to_bytes(which: int, a) -> bytes:
    out = np.zeros(32, dtype=np.uint8)
    lib().orc_to_bytes(which, ptr(np.ascontiguousarray(a, dtype=np.uint64)), ptr(out))
    return out.tobytes()


def from_hex_le(which: int, h: str):
    return from_bytes(which, bytes.fromhex(h))


def to_hex_le(which: int, a) -> str:
    return to_bytes(which, a).hex()


def from_u64(which: int, x: int):
    out = fe()
    lib().orc_from_u64(which, x, ptr(out))
    return out


def binop(name, which, a, b):
    out = fe()
    getattr(lib(), "orc_" + name)(which, ptr(np.ascontiguousarray(a)), ptr(np.ascontiguousarray(b)), ptr(out))
    return out


def mul(w, a, b):
    return binop("mul", w, a, b)


def add(w, a, b):
    return binop("add", w, a, b)


def sub(w, a, b):
    return binop("sub", w, a, b)


def square(w, a):
    out = fe()
    lib().orc_square(w, ptr(np.ascontiguousarray(a)), ptr(out))
    return out


def neg(w, a):
    out = fe()
    lib().orc_neg(w, ptr(np.ascontiguousarray(a)), ptr(out))
    return out


def inverse(w, a):
    out = fe()
    ok = lib().orc_inverse(w, ptr(np.ascontiguousarray(a)), ptr(out))
    return out if ok else None


def pow_(w, a, e_limbs):
    out = fe()
    e = np.ascontiguousarray(e_limbs, dtype=np.uint64)
    lib().orc_pow(w, ptr(np.ascontiguousarray(a)), ptr(e), ptr(out))
    return out


def one(w=FR):
    return from_u64(w, 1)


# ---------------------------------------------------------------- G1
def generator_affine():
    out = np.zeros(8, dtype=np.uint64)
    lib().orc_g1_generator(ptr(out))
    return out


def jac_from_affine(a):
    out = np.zeros(12, dtype=np.uint64)
    lib().orc_g1_from_affine(ptr(np.ascontiguousarray(a)), ptr(out))
    return out


def jac_to_affine(p):
    out = np.zeros(8, dtype=np.uint64)
    lib().orc_g1_to_affine(ptr(np.ascontiguousarray(p)), ptr(out))
    return out


def jac_add(p, q):
    out = np.zeros(12, dtype=np.uint64)
    lib().orc_g1_add(ptr(np.ascontiguousarray(p)), ptr(np.ascontiguousarray(q)), ptr(out))
    return out


def jac_add_affine(p, q):
    out = np.zeros(12, dtype=np.uint64)
    lib().orc_g1_add_affine(ptr(np.ascontiguousarray(p)), ptr(np.ascontiguousarray(q)), ptr(out))
    return out


def jac_double(p):
    out = np.zeros(12, dtype=np.uint64)
    lib().orc_g1_double(ptr(np.ascontiguousarray(p)), ptr(out))
    return out


def jac_mul(p, s):
    out = np.zeros(12, dtype=np.uint64)
    lib().orc_g1_mul(ptr(np.ascontiguousarray(p)), ptr(np.ascontiguousarray(s)), ptr(out))
    return out


def jac_mul_le_bytes(p, k: bytes):
    out = np.zeros(12, dtype=np.uint64)
    kb = np.frombuffer(bytes(k), dtype=np.uint8).copy()
    lib().orc_g1_mul_le_bytes(ptr(np.ascontiguousarray(p)), ptr(kb), ptr(out))
    return out


def affine_to_wire(a) -> bytes:
    out = np.zeros(64, dtype=np.uint8)
    lib().orc_g1_affine_to_wire(ptr(np.ascontiguousarray(a)), ptr(out))
    return out.tobytes()


def affine_from_wire(b: bytes):
    out = np.zeros(8, dtype=np.uint64)
    buf = np.frombuffer(bytes(b), dtype=np.uint8).copy()
    ok = lib().orc_g1_affine_from_wire(ptr(buf), ptr(out))
    return out if ok else None


def jac_to_wire(p) -> bytes:
    return affine_to_wire(jac_to_affine(p))


# ---------------------------------------------------------------- MSM
def window_bits(n: int) -> int:
    return lib().orc_msm_window_bits(n)


def msm(bases_aff: np.ndarray, scalars: np.ndarray, threads: int = 8, workers: int = 1):
    n = scalars.shape[0]
    out = np.zeros(12, dtype=np.uint64)
    lib().orc_msm(ptr(np.ascontiguousarray(bases_aff)), ptr(np.ascontiguousarray(scalars)), n, threads, workers, ptr(out))
    return out


def msm_recode(scalar, w: int) -> np.ndarray:
    """One scalar's signed digits at window width w (msm.rs:180-198)."""
    out = np.zeros(256, dtype=np.int32)
    nd = lib().orc_msm_recode(ptr(np.ascontiguousarray(scalar)), w, ptr(out))
    return out[:nd].copy()


def msm_nonzero_digits(scalars: np.ndarray, w: int = 0) -> int:
    """Nonzero signed digits of the scalars at width w (0: the reference's rule)."""
    s = np.ascontiguousarray(scalars)
    return int(lib().orc_msm_nonzero_digits(ptr(s), s.shape[0], w))


def msm_small_u32(bases_aff, scalars_u32, threads: int = 8, workers: int = 1):
    n = scalars_u32.shape[0]
    out = np.zeros(12, dtype=np.uint64)
    s = np.ascontiguousarray(scalars_u32, dtype=np.uint32)
    lib().orc_msm_small_u32(ptr(np.ascontiguousarray(bases_aff)), ptr(s), n, threads, workers, ptr(out))
    return out


def srs_powers(tau, n: int, workers: int = 8):
    out = np.zeros((n, 8), dtype=np.uint64)
    lib().orc_srs_powers(ptr(np.ascontiguousarray(tau)), n, workers, ptr(out))
    return out


def gen_mul_batch(scalars, workers: int = 8):
    n = scalars.shape[0]
    out = np.zeros((n, 8), dtype=np.uint64)
    lib().orc_g1_gen_mul_batch(ptr(np.ascontiguousarray(scalars)), n, workers, ptr(out))
    return out


# ---------------------------------------------------------------- poly
def bind(table: np.ndarray, r) -> np.ndarray:
    t = np.ascontiguousarray(table.copy())
    lib().orc_bind(ptr(t), t.shape[0], ptr(np.ascontiguousarray(r)))
    return t[: t.shape[0] // 2].copy()


def eq_table(r: np.ndarray) -> np.ndarray:
    n = r.shape[0]
    out = fe(1 << n)
    lib().orc_eq_table(ptr(np.ascontiguousarray(r)), n, ptr(out))
    return out


def eq_eval(r, y):
    out = fe()
    lib().orc_eq_eval(ptr(np.ascontiguousarray(r)), ptr(np.ascontiguousarray(y)), r.shape[0], ptr(out))
    return out


def evaluate(table, point):
    out = fe()
    lib().orc_evaluate(ptr(np.ascontiguousarray(table)), point.shape[0], ptr(np.ascontiguousarray(point)), ptr(out))
    return out


# ---------------------------------------------------------------- transcript
class OrcTranscript(C.Structure):
    _fields_ = [
        ("state", C.c_uint64 * 12),
        ("input", C.c_uint64 * 8),
        ("input_len", C.c_uint64),
        ("output", C.c_uint64 * 8),
        ("output_len", C.c_uint64),
    ]


class Transcript:
    """The reference's duplex transcript (crates/transcript/src/lib.rs)."""

    def __init__(self):
        self.t = OrcTranscript()
        lib().orc_tr_init(C.byref(self.t))

    def ref(self):
        return C.byref(self.t)

    def observe(self, x):
        lib().orc_tr_observe(self.ref(), ptr(np.ascontiguousarray(x)))

    def sample(self):
        out = fe()
        lib().orc_tr_sample(self.ref(), ptr(out))
        return out

    def append_scalars(self, tag, xs):
        xs = np.ascontiguousarray(np.asarray(xs, dtype=np.uint64).reshape(-1, 4))
        lib().orc_tr_append_scalars(self.ref(), tag, ptr(xs), xs.shape[0])

    def append_bytes(self, tag, b: bytes):
        buf = np.frombuffer(bytes(b) + b"\0", dtype=np.uint8).copy()
        lib().orc_tr_append_bytes(self.ref(), tag, ptr(buf), len(b))

    def challenge(self, tag):
        out = fe()
        lib().orc_tr_challenge(self.ref(), tag, ptr(out))
        return out

    def append_g1(self, tag, wires):
        buf = np.frombuffer(b"".join(wires) + b"\0", dtype=np.uint8).copy()
        lib().orc_tr_append_g1(self.ref(), tag, ptr(buf), len(wires))

    def snapshot(self) -> bytes:
        return bytes(self.t)


def mercury_open(srs: np.ndarray, table: np.ndarray, cm_wire: bytes, u: np.ndarray,
                 transcript: "Transcript", workers: int = 8):
    """pcs::open (crates/pcs/src/lib.rs:313-492): returns (v, proof bytes),
    advancing ``transcript``. Raises ValueError with the oracle's code."""
    srs = np.ascontiguousarray(srs, dtype=np.uint64)
    table = np.ascontiguousarray(table, dtype=np.uint64)
    u = np.ascontiguousarray(np.asarray(u, dtype=np.uint64).reshape(-1, 4))
    nv = u.shape[0]
    cm = np.frombuffer(bytes(cm_wire), dtype=np.uint8).copy()
    v = fe()
    proof = np.zeros(704, dtype=np.uint8)
    rc = lib().orc_mercury_open(ptr(srs), srs.shape[0], ptr(table), nv, ptr(cm), ptr(u),
                                transcript.ref(), workers, ptr(v), ptr(proof))
    if rc != 0:
        raise ValueError(f"orc_mercury_open: {rc}")
    return v, proof.tobytes()


def mercury_batch_open_stacked(srs: np.ndarray, stacks, cms_wire, u: np.ndarray, r, transcript: "Transcript",
                               workers: int = 8):
    """batch_open_stacked (crates/pcs/src/lib.rs:701-789); stacks is a list of
    lists of columns ((2^width, 4) arrays). Returns (vs, proof bytes)."""
    srs = np.ascontiguousarray(srs, dtype=np.uint64)
    sizes = np.array([len(s) for s in stacks], dtype=np.uint32)
    flat = [c for s in stacks for c in s]
    cols = np.ascontiguousarray(np.concatenate(flat) if flat else np.zeros((1, 4), dtype=np.uint64), dtype=np.uint64)
    u = np.ascontiguousarray(np.asarray(u, dtype=np.uint64).reshape(-1, 4))
    nr = len(r)
    r = np.ascontiguousarray(np.asarray(r, dtype=np.uint64).reshape(-1, 4)) if nr else np.zeros((1, 4), np.uint64)
    cms = np.frombuffer(b"".join(bytes(c) for c in cms_wire) + b"\0", dtype=np.uint8).copy()
    vs = fe(max(len(stacks), 1))
    proof = np.zeros(704, dtype=np.uint8)
    rc = lib().orc_mercury_batch_open_stacked(ptr(srs), srs.shape[0], ptr(cols), ptr(sizes) if len(sizes) else None,
                                              len(stacks), u.shape[0], ptr(cms), ptr(u), ptr(r), nr,
                                              transcript.ref(), workers, ptr(vs), ptr(proof))
    if rc != 0:
        raise ValueError(f"orc_mercury_batch_open_stacked: {rc}")
    return vs[:len(stacks)], proof.tobytes()


def poseidon2_permute(state: np.ndarray) -> np.ndarray:
    s = np.ascontiguousarray(state.copy().reshape(3, 4))
    lib().orc_poseidon2_permute(ptr(s))
    return s


def g1_limbs(wire: bytes) -> np.ndarray:
    out = fe(4)
    buf = np.frombuffer(bytes(wire), dtype=np.uint8).copy()
    lib().orc_g1_limbs(ptr(buf), ptr(out))
    return out


# ---------------------------------------------------------------- sumcheck
class OrcTerm(C.Structure):
    _fields_ = [("coef", C.c_uint64 * 4), ("f", C.c_int32 * 3)]


def make_terms(terms):
    """terms: list of (coef_limbs, [factor indices]) with 1..3 factors."""
    arr = (OrcTerm * max(1, len(terms)))()
    for i, (coef, fs) in enumerate(terms):
        for k in range(4):
            arr[i].coef[k] = int(coef[k])
        fl = list(fs) + [-1] * (3 - len(fs))
        for k in range(3):
            arr[i].f[k] = fl[k]
    return arr


def interpolate(vals: np.ndarray) -> np.ndarray:
    d = vals.shape[0] - 1
    out = fe(d + 1)
    lib().orc_interpolate(ptr(np.ascontiguousarray(vals)), d, ptr(out))
    return out


def _table_ptrs(tables):
    arr = (C.c_void_p * len(tables))()
    for i, t in enumerate(tables):
        arr[i] = t.ctypes.data
    return arr


def gate_degree(terms) -> int:
    mx = max([len(fs) for _, fs in terms] + [1])
    return max(3, mx + 1)


def prove_sumcheck(eq_point, terms, tables, transcript: Transcript, workers: int = 1):
    """gkr::prove_sumcheck with a gate summand. Tables are copied (the copies
    are left bound); returns (rounds[n, d+1, 4], point[n, 4])."""
    n = eq_point.shape[0]
    d = gate_degree(terms)
    work = [np.ascontiguousarray(t.copy()) for t in tables]
    rounds = np.zeros((n, d + 1, 4), dtype=np.uint64)
    point = fe(n)
    tp = make_terms(terms)
    lib().orc_prove_sumcheck(ptr(np.ascontiguousarray(eq_point)), n, tp, len(terms), _table_ptrs(work), len(work),
                             transcript.ref(), workers, ptr(rounds), ptr(point))
    return rounds, point, [w[:1].copy() for w in work]


def prove_zerocheck(terms, tables, transcript: Transcript):
    n = int(tables[0].shape[0]).bit_length() - 1
    d = gate_degree(terms)
    work = [np.ascontiguousarray(t.copy()) for t in tables]
    rounds = np.zeros((n, d + 1, 4), dtype=np.uint64)
    finals = fe(len(tables))
    tp = make_terms(terms)
    lib().orc_prove_zerocheck(tp, len(terms), _table_ptrs(work), len(work), n, transcript.ref(), ptr(rounds), ptr(finals))
    return rounds, finals


def witness_digest(tables):
    n = int(tables[0].shape[0]).bit_length() - 1
    work = [np.ascontiguousarray(t) for t in tables]
    out = fe()
    lib().orc_witness_digest(_table_ptrs(work), len(work), n, ptr(out))
    return out


def combine(cols: np.ndarray, weights: np.ndarray, outs: int) -> np.ndarray:
    """cols (k, n, 4), weights (outs * k, 4) -> (outs, n, 4)."""
    k, n = cols.shape[0], cols.shape[1]
    out = np.zeros((outs, n, 4), dtype=np.uint64)
    lib().orc_combine(ptr(np.ascontiguousarray(cols)), n, k, ptr(np.ascontiguousarray(weights)), outs, ptr(out))
    return out


def axpy(a, x: np.ndarray, y: np.ndarray) -> None:
    """y += a * x in place (Fr, (n, 4) arrays)."""
    assert y.flags["C_CONTIGUOUS"] and x.shape == y.shape
    lib().orc_axpy(x.shape[0], ptr(np.ascontiguousarray(a)), ptr(np.ascontiguousarray(x)), ptr(y))


def product_claim(r: np.ndarray, tables) -> np.ndarray:
    """sum_x eq(r, x) prod_o t_o(x) over the hypercube of r's length."""
    work = [np.ascontiguousarray(t) for t in tables]
    out = fe()
    lib().orc_product_claim(ptr(np.ascontiguousarray(r)), r.shape[0], _table_ptrs(work), len(work), ptr(out))
    return out


def shard_script(srs_aff, cols: np.ndarray, degree: int, transcript: "Transcript", workers: int = 8):
    """The shard composition (commit k columns, absorb as one COMMITMENT
    message, rho = MERCURY_BATCH, f = degree - 1 combinations with weights
    rho^(o k + c), eq point = num_vars SUMCHECK_CHALLENGE draws, the rounds of
    eq * t_0 * ... * t_(f-1), finals under SUMCHECK_FINAL_EVALS), built from
    the restated primitives. Returns (commitment wires, rounds, finals)."""
    k, n = cols.shape[0], cols.shape[1]
    nv = n.bit_length() - 1
    cms = [jac_to_wire(msm(srs_aff[:n], cols[c], threads=8, workers=workers)) for c in range(k)]
    transcript.append_g1(TAG_COMMITMENT, cms)
    rho = transcript.challenge(TAG_MERCURY_BATCH)
    f = degree - 1
    w = fe(f * k)
    acc = from_u64(FR, 1)
    for i in range(f * k):
        w[i] = acc
        acc = mul(FR, acc, rho)
    tables = combine(cols, w, f)
    r = np.stack([transcript.challenge(TAG_SUMCHECK_CHALLENGE) for _ in range(nv)])
    terms = [(from_u64(FR, 1), list(range(f)))]
    rounds, point, finals = prove_sumcheck(r, terms, [tables[o] for o in range(f)], transcript, workers=workers)
    fin = np.stack([x[0] for x in finals])
    transcript.append_scalars(TAG_SUMCHECK_FINAL_EVALS, fin)
    return cms, rounds, fin


# ---------------------------------------------------------------- inputs
class SplitMix:
    def __init__(self, seed: int):
        self.s = np.array([seed], dtype=np.uint64)

    def next_u64(self) -> int:
        return int(lib().orc_splitmix_next(ptr(self.s)))

    def next_fr(self):
        out = fe()
        lib().orc_splitmix_next_fr(ptr(self.s), ptr(out))
        return out


def chacha_block(key32: bytes, counter: int, nonce12: bytes) -> bytes:
    k = np.frombuffer(key32, dtype=np.uint32).copy()
    nn = np.frombuffer(nonce12, dtype=np.uint32).copy()
    out = np.zeros(64, dtype=np.uint8)
    lib().orc_chacha_block(ptr(k), counter, ptr(nn), ptr(out))
    return out.tobytes()


def fill_uniform(seed: int, n: int) -> np.ndarray:
    out = fe(n)
    lib().orc_fill_uniform(seed, ptr(out), n)
    return out


def fill_ctr(seed: int, n: int, kind: str = "witness") -> np.ndarray:
    """Counter-mode ChaCha20 draws (the benchmark inputs)."""
    out = fe(n)
    lib().orc_fill_ctr(seed, {"uniform": 0, "witness": 1}[kind], ptr(out), n)
    return out


def fill_witness(seed: int, n: int) -> np.ndarray:
    out = fe(n)
    lib().orc_fill_witness(seed, ptr(out), n)
    return out


# ---------------------------------------------------------------- GKR layer (oracle/gkr.c)
GATE_LINEAR, GATE_PRODUCT, GATE_MASK, GATE_AFFINE_PRODUCT, GATE_TREE_PRODUCT, GATE_QUADRATIC, GATE_TREE_CROSS = range(7)
VIRT_ROW_INDEX, VIRT_RAM_LIVE, VIRT_RANGE19, VIRT_RANGE16, VIRT_XOR8_A, VIRT_XOR8_B, VIRT_XOR8_OUT = range(7)
SRC_LOWER, SRC_UPPER, SRC_VIRTUAL, SRC_CACHED = range(4)


class OrcGkrList(C.Structure):
    _fields_ = [
        ("gates", C.c_void_p),
        ("ncached", C.c_uint32),
        ("nproducing", C.c_uint32),
        ("nenforcing", C.c_uint32),
        ("sources", C.c_void_p),
        ("nsources", C.c_uint32),
        ("coeffs", C.c_void_p),
        ("ncoeffs", C.c_uint32),
        ("virtual_kinds", C.c_void_p),
        ("nvirtual", C.c_uint32),
        ("halving", C.c_uint32),
    ]


class GkrFlat:
    """A gate list flattened as ResolvedList::new resolves it: gates (N, 5)
    uint32 rows (shape, src_begin, src_end, coef_begin, split), sources
    uint32 words (kind << 24 | index), coeffs (M, 4) Montgomery limbs, the
    virtual kinds, and the counts."""

    def __init__(self, gates, ncached, nproducing, nenforcing, sources, coeffs, kinds, halving):
        self.gates = np.ascontiguousarray(np.asarray(gates, dtype=np.uint32).reshape(-1, 5))
        self.sources = np.ascontiguousarray(np.asarray(sources, dtype=np.uint32).reshape(-1))
        self.coeffs = np.ascontiguousarray(np.asarray(coeffs, dtype=np.uint64).reshape(-1, 4))
        self.kinds = np.ascontiguousarray(np.asarray(kinds, dtype=np.uint32).reshape(-1))
        self.ncached, self.nproducing, self.nenforcing = ncached, nproducing, nenforcing
        self.halving = bool(halving)
        self._s = OrcGkrList(
            self.gates.ctypes.data, ncached, nproducing, nenforcing, self.sources.ctypes.data,
            self.sources.shape[0], self.coeffs.ctypes.data, self.coeffs.shape[0], self.kinds.ctypes.data,
            self.kinds.shape[0], 1 if halving else 0)

    def ref(self):
        return C.byref(self._s)

    @property
    def nweighted(self):
        return self.nproducing + self.nenforcing


def _arr(x, n=None):
    a = np.ascontiguousarray(np.asarray(x, dtype=np.uint64).reshape(-1, 4))
    if a.shape[0] == 0:
        a = fe(1)
    return a


def gkr_eval_gate(shape: int, coeffs, split: int, values):
    co = _arr(coeffs)
    v = _arr(values)
    nv = 0 if values is None else np.asarray(values, dtype=np.uint64).reshape(-1, 4).shape[0]
    out = fe()
    lib().orc_gkr_eval_gate(shape, ptr(co), split, ptr(v), nv, ptr(out))
    return out


def gkr_virtual_at_row(kind: int, row: int):
    out = fe()
    lib().orc_gkr_virtual_at_row(kind, row, ptr(out))
    return out


def gkr_virtual_at_point(kind: int, point):
    p = _arr(point)
    n = np.asarray(point, dtype=np.uint64).reshape(-1, 4).shape[0]
    out = fe()
    lib().orc_gkr_virtual_at_point(kind, ptr(p), n, ptr(out))
    return out


def _n(x):
    return 0 if x is None else np.asarray(x, dtype=np.uint64).reshape(-1, 4).shape[0]


def gkr_gate_values(flat: GkrFlat, lower, upper=None, virtuals=None):
    out = fe(max(flat.nweighted, 1))
    lib().orc_gkr_gate_values(flat.ref(), ptr(_arr(lower)), _n(lower), ptr(_arr(upper if upper is not None else [])),
                              _n(upper), ptr(_arr(virtuals if virtuals is not None else [])), ptr(out))
    return out[:flat.nweighted]


def gkr_summand(flat: GkrFlat, weights, lower, upper=None, virtuals=None):
    out = fe()
    lib().orc_gkr_summand(flat.ref(), ptr(_arr(weights)), ptr(_arr(lower)), _n(lower),
                          ptr(_arr(upper if upper is not None else [])), _n(upper),
                          ptr(_arr(virtuals if virtuals is not None else [])), ptr(out))
    return out


def gkr_forward_layer(flat: GkrFlat, cols, num_vars: int):
    """gkr::produce for one list: one output column per producing gate."""
    cols = [np.ascontiguousarray(c) for c in cols]
    rows = (1 << num_vars) // (2 if flat.halving else 1)
    outs = [fe(rows) for _ in range(flat.nproducing)]
    lib().orc_gkr_forward_layer(flat.ref(), _table_ptrs(cols), len(cols), num_vars, _table_ptrs(outs))
    return outs


def gkr_prove_sumcheck(flat: GkrFlat, weights, eq_point, lower, upper, transcript: Transcript):
    """gkr::prove_sumcheck for the weighted list. Tables are copied (the
    copies are left bound); returns (rounds[n, 4, 4], point[n, 4], finals:
    the bound lower then upper values)."""
    n = _n(eq_point)
    lo = [np.ascontiguousarray(t.copy()) for t in lower]
    up = [np.ascontiguousarray(t.copy()) for t in (upper or [])]
    rounds = np.zeros((max(n, 1), 4, 4), dtype=np.uint64)
    point = fe(max(n, 1))
    lib().orc_gkr_prove_sumcheck(flat.ref(), ptr(_arr(weights)), ptr(_arr(eq_point)), n, _table_ptrs(lo),
                                 _table_ptrs(up) if up else None, len(lo), transcript.ref(), ptr(rounds), ptr(point))
    return rounds[:n], point[:n], np.stack([t[0] for t in lo + up])
