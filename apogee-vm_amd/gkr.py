"""The general GKR layer through the C ABI.

Mirrors the reference's types and entry points for one gate list:
- the gate shapes, addresses and coefficients of ``constraints``
  (crates/constraints/src/lib.rs:215-430) and a circuit artifact's layer
  lists (``CircuitArtifact``, ``LayerSpec``, crates/constraints/src/lib.rs:611);
- ``ResolvedList`` (crates/gkr-verify/src/lib.rs:300-395): a list flattened to
  the C ABI's ``apg_gkr_list``, every operand resolved once to the index its
  value is read at;
- ``forward_layer`` / ``forward`` (gkr::produce / gkr::forward,
  crates/gkr/src/lib.rs:279-336) and ``prove_sumcheck``
  (gkr::prove_sumcheck with a ``LayerSummand``, crates/gkr/src/lib.rs:687-899),
  computed by the device.

Field elements are numpy ``uint64`` arrays of shape ``(..., 4)`` holding the
reference's Montgomery limbs, as everywhere in the package.
"""
from __future__ import annotations

import ctypes as C
import enum
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple, Union

import numpy as np

from . import ROUND_FN, Context, DeviceBuffer, Transcript, _check, _fe_array, _p, default_context, lib


class VirtualKind(enum.IntEnum):
    """constraints::VirtualKind, in wire order (crates/constraints/src/lib.rs:215-250)."""

    RowIndex = 0
    RamLive = 1
    Range19 = 2
    Range16 = 3
    Xor8A = 4
    Xor8B = 5
    Xor8Out = 6


# ---- addresses (constraints::PolyAddress, crates/constraints/src/lib.rs:253-268)
@dataclass(frozen=True)
class Memory:
    i: int


@dataclass(frozen=True)
class Witness:
    i: int


@dataclass(frozen=True)
class Setup:
    i: int


@dataclass(frozen=True)
class Virtual:
    kind: VirtualKind


@dataclass(frozen=True)
class Inner:
    layer: int
    offset: int


@dataclass(frozen=True)
class Cached:
    layer: int
    offset: int


PolyAddress = Union[Memory, Witness, Setup, Virtual, Inner, Cached]


# ---- coefficients (constraints::Coeff, crates/constraints/src/lib.rs:299-305)
@dataclass(frozen=True)
class Literal:
    value: tuple  # four Montgomery limbs

    @staticmethod
    def of(fe) -> "Literal":
        return Literal(tuple(int(x) for x in np.asarray(fe, dtype=np.uint64).reshape(4)))


@dataclass(frozen=True)
class Challenge:
    slot: int


Coeff = Union[Literal, Challenge]


# ---- gate shapes (constraints::GateDef, crates/constraints/src/lib.rs:310-430)
@dataclass
class Linear:
    """Σ c_i·x_i + c_0."""

    terms: List[Tuple[Coeff, PolyAddress]]
    constant: Coeff
    TAG = 0

    def operands(self):
        return [a for _, a in self.terms]

    def coefficients(self):
        return [c for c, _ in self.terms] + [self.constant]

    split = 0


@dataclass
class Product:
    """c·x·y."""

    coeff: Coeff
    left: PolyAddress
    right: PolyAddress
    TAG = 1
    split = 0

    def operands(self):
        return [self.left, self.right]

    def coefficients(self):
        return [self.coeff]


@dataclass
class MaskIntoIdentity:
    """x·m + (1 − m)."""

    input: PolyAddress
    mask: PolyAddress
    TAG = 2
    split = 0

    def operands(self):
        return [self.input, self.mask]

    def coefficients(self):
        return []


@dataclass
class AffineProduct:
    """(Σ a_i·x_i + a_0)·(Σ b_j·y_j + b_0)."""

    left: List[Tuple[Coeff, PolyAddress]]
    left_constant: Coeff
    right: List[Tuple[Coeff, PolyAddress]]
    right_constant: Coeff
    TAG = 3

    @property
    def split(self):
        return len(self.left)

    def operands(self):
        return [a for _, a in self.left] + [a for _, a in self.right]

    def coefficients(self):
        return [c for c, _ in self.left] + [self.left_constant] + [c for c, _ in self.right] + [self.right_constant]


@dataclass
class TreeProduct:
    """x(·,0)·x(·,1), in a halving list."""

    input: PolyAddress
    TAG = 4
    split = 0

    def operands(self):
        return [self.input]

    def coefficients(self):
        return []


@dataclass
class Quadratic:
    """c_0 + Σ a_i·x_i + Σ b_j·y_j·z_j."""

    constant: Coeff
    linear: List[Tuple[Coeff, PolyAddress]]
    products: List[Tuple[Coeff, PolyAddress, PolyAddress]]
    TAG = 5

    @property
    def split(self):
        return len(self.linear)

    def operands(self):
        out = [a for _, a in self.linear]
        for _, y, z in self.products:
            out += [y, z]
        return out

    def coefficients(self):
        return [self.constant] + [c for c, _ in self.linear] + [b for b, _, _ in self.products]


@dataclass
class TreeCross:
    """p(·,0)·q(·,1) + p(·,1)·q(·,0), in a halving list."""

    left: PolyAddress
    right: PolyAddress
    TAG = 6
    split = 0

    def operands(self):
        return [self.left, self.right]

    def coefficients(self):
        return []


GateDef = Union[Linear, Product, MaskIntoIdentity, AffineProduct, TreeProduct, Quadratic, TreeCross]


@dataclass
class LayerSpec:
    """One gate list: it reads layer k and writes layer k + 1
    (crates/constraints/src/lib.rs LayerSpec). Entries are (name, address,
    gate) for cached and producing lists, (name, gate) for enforcing."""

    halving: bool
    num_vars: int
    width: int
    cached: list = field(default_factory=list)
    producing: list = field(default_factory=list)
    enforcing: list = field(default_factory=list)


@dataclass
class CircuitArtifact:
    """The layered part of constraints::CircuitArtifact: the committed column
    names (layout order memory, witness, setup), the virtual tables and the
    gate lists."""

    memory: List[str]
    witness: List[str]
    setup: List[str]
    virtuals: List[Tuple[VirtualKind, str]]
    layers: List[LayerSpec]

    def committed(self) -> int:
        return len(self.memory) + len(self.witness) + len(self.setup)

    def depth(self) -> int:
        return len(self.layers)


class GkrError(Exception):
    pass


SRC_LOWER, SRC_UPPER, SRC_VIRTUAL, SRC_CACHED = range(4)


class ApgGkrGate(C.Structure):
    _fields_ = [(f, C.c_uint32) for f in ("shape", "src_begin", "src_end", "coef_begin", "split")]


class ApgGkrList(C.Structure):
    _fields_ = [
        ("gates", C.c_void_p),
        ("ncached", C.c_uint32),
        ("nproducing", C.c_uint32),
        ("nenforcing", C.c_uint32),
        ("sources", C.c_void_p),
        ("nsources", C.c_uint32),
        ("coeffs_mont", C.c_void_p),
        ("ncoeffs", C.c_uint32),
        ("virtual_kinds", C.c_void_p),
        ("nvirtual", C.c_uint32),
        ("halving", C.c_uint32),
    ]


def _column_index(artifact: CircuitArtifact, k: int, op) -> int:
    """crates/gkr-verify/src/lib.rs:541-550."""
    m, w = len(artifact.memory), len(artifact.witness)
    if isinstance(op, Memory):
        return op.i
    if isinstance(op, Witness):
        return m + op.i
    if isinstance(op, Setup):
        return m + w + op.i
    if isinstance(op, Inner):
        return op.offset
    raise GkrError(f"gate list {k} cannot read {op} as a column")


class ResolvedList:
    """Gate list ``k`` of ``artifact`` with every operand resolved once
    (crates/gkr-verify/src/lib.rs:330-395): a halving gate reads each operand
    at both children, column x giving lower[x] then upper[x]; a row-wise
    list's gates are its cached entries, producing gates and enforcing gates,
    a virtual operand resolving to the artifact's table index and a cached one
    to its offset. Challenge coefficients take their values from
    ``challenges`` (slot -> Fr)."""

    def __init__(self, artifact: CircuitArtifact, k: int, challenges: Optional[Dict[int, np.ndarray]] = None):
        challenges = challenges or {}
        spec = artifact.layers[k]
        self.halving = bool(spec.halving)
        gates, sources, coeffs = [], [], []

        def coef_value(c):
            if isinstance(c, Literal):
                return np.array(c.value, dtype=np.uint64)
            if c.slot not in challenges:
                raise GkrError(f"external challenge slot {c.slot} has no value")
            return np.asarray(challenges[c.slot], dtype=np.uint64).reshape(4)

        def push(gate, srcs):
            cs = gate.coefficients()
            gates.append((gate.TAG, len(sources), len(sources) + len(srcs), len(coeffs), gate.split))
            sources.extend(srcs)
            coeffs.extend(coef_value(c) for c in cs)

        if self.halving:
            for entry in spec.producing:
                gate = entry[-1]
                srcs = []
                for op in gate.operands():
                    x = _column_index(artifact, k, op)
                    srcs += [(SRC_LOWER << 24) | x, (SRC_UPPER << 24) | x]
                push(gate, srcs)
            self.ncached, self.nenforcing = 0, 0
        else:
            kinds = [v for v, _ in artifact.virtuals]

            def resolve(op):
                if isinstance(op, Virtual):
                    if op.kind not in kinds:
                        raise GkrError(f"the artifact lists no virtual table {op.kind!r}")
                    return (SRC_VIRTUAL << 24) | kinds.index(op.kind)
                if isinstance(op, Cached):
                    return (SRC_CACHED << 24) | op.offset
                return (SRC_LOWER << 24) | _column_index(artifact, k, op)

            for entry in list(spec.cached) + list(spec.producing) + [e for e in spec.enforcing]:
                gate = entry[-1]
                push(gate, [resolve(op) for op in gate.operands()])
            self.ncached, self.nenforcing = len(spec.cached), len(spec.enforcing)
        self.nproducing = len(spec.producing)
        self.gates = np.array(gates, dtype=np.uint32).reshape(-1, 5)
        self.sources = np.array(sources, dtype=np.uint32).reshape(-1)
        self.coeffs = np.ascontiguousarray(np.array(coeffs, dtype=np.uint64).reshape(-1, 4))
        self.kinds = np.array([] if self.halving else [int(v) for v, _ in artifact.virtuals], dtype=np.uint32)
        self._struct = None

    @classmethod
    def from_arrays(cls, gates, ncached, nproducing, nenforcing, sources, coeffs, kinds, halving) -> "ResolvedList":
        """A list already flattened (the C ABI's apg_gkr_list fields)."""
        self = cls.__new__(cls)
        self.halving = bool(halving)
        self.gates = np.ascontiguousarray(np.asarray(gates, dtype=np.uint32).reshape(-1, 5))
        self.sources = np.ascontiguousarray(np.asarray(sources, dtype=np.uint32).reshape(-1))
        self.coeffs = np.ascontiguousarray(np.asarray(coeffs, dtype=np.uint64).reshape(-1, 4))
        self.kinds = np.ascontiguousarray(np.asarray(kinds, dtype=np.uint32).reshape(-1))
        self.ncached, self.nproducing, self.nenforcing = ncached, nproducing, nenforcing
        self._struct = None
        return self

    def producing(self) -> int:
        return self.nproducing

    def enforcing(self) -> int:
        return self.nenforcing

    def arrays(self) -> dict:
        return dict(gates=self.gates, ncached=self.ncached, nproducing=self.nproducing,
                    nenforcing=self.nenforcing, sources=self.sources, coeffs=self.coeffs, kinds=self.kinds,
                    halving=self.halving)

    def c_list(self) -> ApgGkrList:
        if self._struct is None:
            for name in ("gates", "sources", "coeffs", "kinds"):
                setattr(self, name, np.ascontiguousarray(getattr(self, name)))
            self._struct = ApgGkrList(self.gates.ctypes.data, self.ncached, self.nproducing, self.nenforcing,
                                      self.sources.ctypes.data, self.sources.shape[0], self.coeffs.ctypes.data,
                                      self.coeffs.shape[0], self.kinds.ctypes.data, self.kinds.shape[0],
                                      1 if self.halving else 0)
        return self._struct


def _ptrs(arrs):
    return (C.c_void_p * max(len(arrs), 1))(*[a.ctypes.data for a in arrs])


def _tables(tables) -> List[np.ndarray]:
    return [np.ascontiguousarray(np.asarray(t, dtype=np.uint64).reshape(-1, 4)) for t in tables]


def forward_layer(resolved: ResolvedList, columns: Sequence, num_vars: int,
                  ctx: Optional[Context] = None) -> List[np.ndarray]:
    """gkr::produce for one list (crates/gkr/src/lib.rs:300-336): the producing
    gates at every row of layer k (a halving list: at every row of the half,
    reading rows y and y + half)."""
    ctx = ctx or default_context()
    cols = _tables(columns)
    if any(c.shape[0] != 1 << num_vars for c in cols):
        raise GkrError("forward_layer: every column has 2^num_vars rows")
    if resolved.halving and num_vars < 1:
        raise GkrError("forward_layer: a halving list needs num_vars >= 1")
    rows = (1 << num_vars) // (2 if resolved.halving else 1)
    outs = [np.zeros((rows, 4), dtype=np.uint64) for _ in range(resolved.nproducing)]
    lst = resolved.c_list()
    _check(lib().apg_gkr_forward_layer(ctx.h, C.byref(lst), _ptrs(cols), len(cols), num_vars, _ptrs(outs)))
    return outs


def forward_layer_device(resolved: ResolvedList, columns: Sequence[DeviceBuffer], num_vars: int,
                         out: Sequence[DeviceBuffer], ctx: Optional[Context] = None) -> None:
    """forward_layer on device buffers (2^num_vars elements each in, one
    output buffer per producing gate)."""
    ctx = ctx or default_context()
    lst = resolved.c_list()
    cols = (C.c_void_p * max(len(columns), 1))(*[b.ptr for b in columns])
    outs = (C.c_void_p * max(len(out), 1))(*[b.ptr for b in out])
    _check(lib().apg_gkr_forward_layer_dev(ctx.h, C.byref(lst), cols, len(columns), num_vars, outs))


def forward(artifact: CircuitArtifact, base: Sequence, challenges: Optional[Dict[int, np.ndarray]] = None,
            ctx: Optional[Context] = None) -> List[List[np.ndarray]]:
    """gkr::forward (crates/gkr/src/lib.rs:279-296): every layer from the
    committed columns ``base`` (layout order), list by list. Returns
    [layer 1 columns, ..., layer N columns]."""
    layers: List[List[np.ndarray]] = []
    cols = _tables(base)
    nv = cols[0].shape[0].bit_length() - 1 if cols else 0
    for k in range(artifact.depth()):
        r = ResolvedList(artifact, k, challenges)
        out = forward_layer(r, cols, nv, ctx)
        if artifact.layers[k].halving:
            nv -= 1
        layers.append(out)
        cols = out
    return layers


def prove_sumcheck(eq_point, resolved: ResolvedList, weights, lower: Sequence, upper: Optional[Sequence],
                   transcript: Transcript, ctx: Optional[Context] = None):
    """gkr::prove_sumcheck (crates/gkr/src/lib.rs:792-899) for the summand
    Σ_j weights[j]·gate_j of the list (producing, then enforcing gates):
    ``lower`` are the layer's columns (child 0 of each for a halving list)
    and ``upper`` child 1 of each (halving lists only), all with the eq
    point's variable count. Each round is absorbed as SUMCHECK_ROUND and
    followed by a SUMCHECK_CHALLENGE on ``transcript``. Returns (rounds[n, 4],
    point[n], finals: the tables bound to the point, lower then upper)."""
    ctx = ctx or default_context()
    lo = _tables(lower)
    up = _tables(upper) if upper is not None else []
    if resolved.halving and len(up) != len(lo):
        raise GkrError("prove_sumcheck: a halving list needs both children of every column")
    eqp = _fe_array(eq_point) if len(eq_point) else np.zeros((1, 4), dtype=np.uint64)
    n = 0 if not len(eq_point) else eqp.shape[0]
    if any(t.shape[0] != 1 << n for t in lo + up):
        raise GkrError("prove_sumcheck: every table has the eq point's variable count")
    w = np.ascontiguousarray(_fe_array(weights)) if len(weights) else np.zeros((1, 4), dtype=np.uint64)
    if (0 if not len(weights) else w.shape[0]) != resolved.nproducing + resolved.nenforcing:
        raise GkrError("prove_sumcheck: one weight per producing and enforcing gate")
    rounds = np.zeros((max(n, 1), 4, 4), dtype=np.uint64)
    point = np.zeros((max(n, 1), 4), dtype=np.uint64)
    finals = np.zeros((max(len(lo) + len(up), 1), 4), dtype=np.uint64)
    lst = resolved.c_list()
    cb = C.cast(lib().apg_transcript_round, ROUND_FN)
    _check(lib().apg_gkr_prove_sumcheck(ctx.h, C.byref(lst), _p(w), _p(eqp), n, _ptrs(lo),
                                        _ptrs(up) if up else None, len(lo), cb, transcript.h, _p(rounds),
                                        _p(point), _p(finals)))
    return rounds[:n], point[:n], finals[:len(lo) + len(up)]


def prove_sumcheck_device(eq_point, resolved: ResolvedList, weights, lower: Sequence[DeviceBuffer],
                          upper: Optional[Sequence[DeviceBuffer]], transcript: Transcript,
                          ctx: Optional[Context] = None):
    """prove_sumcheck on device-resident tables (read, not modified)."""
    ctx = ctx or default_context()
    up = list(upper) if upper is not None else []
    eqp = _fe_array(eq_point) if len(eq_point) else np.zeros((1, 4), dtype=np.uint64)
    n = 0 if not len(eq_point) else eqp.shape[0]
    w = np.ascontiguousarray(_fe_array(weights)) if len(weights) else np.zeros((1, 4), dtype=np.uint64)
    rounds = np.zeros((max(n, 1), 4, 4), dtype=np.uint64)
    point = np.zeros((max(n, 1), 4), dtype=np.uint64)
    finals = np.zeros((max(len(lower) + len(up), 1), 4), dtype=np.uint64)
    lst = resolved.c_list()
    cb = C.cast(lib().apg_transcript_round, ROUND_FN)
    lp = (C.c_void_p * max(len(lower), 1))(*[b.ptr for b in lower])
    upp = (C.c_void_p * max(len(up), 1))(*[b.ptr for b in up]) if up else None
    _check(lib().apg_gkr_prove_sumcheck_dev(ctx.h, C.byref(lst), _p(w), _p(eqp), n, lp, upp, len(lower), cb,
                                            transcript.h, _p(rounds), _p(point), _p(finals)))
    return rounds[:n], point[:n], finals[:len(lower) + len(up)]
