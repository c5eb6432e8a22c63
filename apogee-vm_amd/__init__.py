"""apogee-vm_amd: the MI355X (gfx950) per-shard proving path of apogee-vm.

Host-side mirror of the reference's interfaces for this path, over the C ABI of
``libapogee_gpu.so`` (include/apogee_gpu.h):

- ``msm`` / ``msm_small_u32``          crates/curve/src/msm.rs:48 / :79
- ``Srs``, ``commit``, ``commit_batch``, ``kzg_commit``
                                        crates/srs/src/lib.rs:96, crates/pcs/src/lib.rs:237,
                                        crates/prover/src/lib.rs:444, crates/srs/src/kzg.rs:20
- ``Gate`` / ``GateTerm``, ``prove_zerocheck``
                                        crates/sumcheck/src/lib.rs:88-151, :359
- ``prove_sumcheck``                    crates/gkr/src/lib.rs:792
- ``Transcript``                        crates/transcript/src/lib.rs:197-346
- ``prove_shard``                       commit + batched sumcheck of one shard

Field elements are numpy ``uint64`` arrays of shape ``(..., 4)`` holding the
reference's in-memory Montgomery limbs (``Fr::to_memory_bytes``); affine G1
points are 64 bytes, Montgomery ``x || y``, all-zero for infinity, or the
canonical wire form where a function says so. There is no CPU fallback: the
library must be built (``python apogee-vm_amd/build.py``) and a GPU present.
"""
from __future__ import annotations

import atexit
import ctypes as C
import enum
import os
import weakref
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# APG_LIB selects another build of the library (A/B timing of variants)
LIB_PATH = os.environ.get("APG_LIB") or os.path.join(HERE, "libapogee_gpu.so")

# Live native handles, released at interpreter exit in dependency order
# (engines, then buffers and SRS tables, then the contexts they belong to)
# rather than in whatever order the interpreter tears objects down.
_live = {kind: weakref.WeakSet() for kind in ("engine", "buffer", "srs", "ctx")}


@atexit.register
def _release_all():
    for kind, method in (("engine", "close"), ("buffer", "free"), ("srs", "release"), ("ctx", "close")):
        for obj in list(_live[kind]):
            try:
                getattr(obj, method)()
            except Exception:
                pass

# crates/constants/src/lib.rs:579-770
TAG_COMMITMENT = 3
TAG_SUMCHECK_ROUND = 4
TAG_SUMCHECK_CHALLENGE = 5
TAG_EVALUATION_CLAIM = 6
TAG_WITNESS_DIGEST = 8
TAG_SUMCHECK_FINAL_EVALS = 9
TAG_MERCURY_BATCH = 17

# crates/pcs-verify/src/lib.rs:89-103
MAX_NUM_VARS = 2 * (28 - 1)

_lib = None


class ApgError(RuntimeError):
    """A negative return code of the C ABI, with apg_last_error()'s message."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"apogee_gpu error {code}: {msg}")
        self.code = code


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is not built; run `python apogee-vm_amd/build.py` (there is no CPU fallback)"
            )
        L = C.CDLL(LIB_PATH)
        _declare(L)
        _lib = L
    return _lib


VP = C.c_void_p
SZ = C.c_size_t
U64 = C.c_uint64
I32 = C.c_int


class ApgTerm(C.Structure):
    _fields_ = [("coef_mont", C.c_uint8 * 32), ("factor", C.c_int32 * 3)]


ROUND_FN = C.CFUNCTYPE(C.c_int, VP, VP, SZ, VP)


def _declare(L):
    def d(name, args, res=I32):
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res

    d("apg_last_error", [], C.c_char_p)
    d("apg_version", [])
    d("apg_ctx_create", [I32, C.POINTER(VP)])
    d("apg_ctx_destroy", [VP], None)
    d("apg_ctx_synchronize", [VP])
    d("apg_dev_alloc", [VP, SZ, C.POINTER(VP)])
    d("apg_dev_free", [VP, VP])
    d("apg_dev_upload", [VP, VP, VP, SZ])
    d("apg_dev_download", [VP, VP, VP, SZ])
    d("apg_dev_copy", [VP, VP, VP, SZ])
    d("apg_srs_upload", [VP, VP, SZ, C.POINTER(VP)])
    d("apg_srs_generate", [VP, VP, SZ, C.POINTER(VP)])
    d("apg_srs_load", [VP, C.c_char_p, C.POINTER(VP), VP, C.POINTER(I32), C.POINTER(U64)])
    d("apg_srs_archive_check", [C.c_char_p, C.POINTER(C.c_uint32), VP, C.POINTER(I32), C.POINTER(U64)])
    d("apg_srs_save", [VP, VP, VP, C.c_char_p])
    d("apg_srs_upload_wire", [VP, VP, SZ, C.POINTER(VP), C.POINTER(U64)])
    d("apg_g2_wire_valid", [VP])
    d("apg_g1_batch_to_affine", [VP, VP, SZ, VP])
    d("apg_srs_download", [VP, VP, SZ, SZ, VP])
    d("apg_srs_len", [VP], SZ)
    d("apg_srs_prepare", [VP, VP, SZ])
    d("apg_srs_device_bytes", [VP], SZ)
    d("apg_msm_set_window_bits", [I32, I32])
    d("apg_msm_window_bits", [SZ])
    d("apg_msm_set_slice_rounds", [I32, I32])
    d("apg_msm_slice_rounds", [SZ])
    d("apg_srs_release", [VP], None)
    d("apg_msm_g1_fr", [VP, VP, VP, SZ, VP])
    d("apg_msm_g1_u32", [VP, VP, VP, SZ, VP])
    d("apg_commit_fr", [VP, VP, VP, SZ, SZ, VP])
    d("apg_commit_u32", [VP, VP, VP, SZ, SZ, VP])
    d("apg_commit_fr_dev", [VP, VP, VP, SZ, SZ, VP])
    d("apg_sumcheck_prove", [VP, VP, SZ, VP, SZ, SZ, VP, ROUND_FN, VP, VP, VP, VP])
    d("apg_sumcheck_prove_dev", [VP, VP, SZ, VP, SZ, SZ, VP, ROUND_FN, VP, VP, VP, VP])
    d("apg_sumcheck_prove_dev_noalloc", [VP, VP, SZ, VP, SZ, SZ, VP, ROUND_FN, VP, VP, VP, VP])
    d("apg_sumcheck_prove_dev_lifted", [VP, VP, SZ, VP, VP, SZ, SZ, VP, ROUND_FN, VP, VP, VP, VP])
    d("apg_debug_imul_peak", [VP, I32, I32, C.POINTER(C.c_double), C.POINTER(C.c_double)])
    d("apg_zerocheck_prove", [VP, VP, SZ, VP, SZ, SZ, VP, VP, VP])
    d("apg_witness_digest", [VP, SZ, SZ, VP])
    d("apg_witness_digest_dev", [VP, VP, SZ, SZ, VP])
    d("apg_prove_shard_dev", [VP, VP, VP, SZ, SZ, I32, VP, VP, VP, VP])
    d("apg_transcript_new", [C.POINTER(VP)])
    d("apg_transcript_free", [VP], None)
    d("apg_transcript_clone", [VP, C.POINTER(VP)])
    d("apg_transcript_observe", [VP, VP], None)
    d("apg_transcript_sample", [VP, VP], None)
    d("apg_transcript_append_scalars", [VP, U64, VP, SZ], None)
    d("apg_transcript_append_bytes", [VP, U64, VP, SZ], None)
    d("apg_transcript_challenge", [VP, U64, VP], None)
    d("apg_transcript_append_g1", [VP, U64, VP, SZ], None)
    d("apg_transcript_round", [VP, VP, SZ, VP])
    d("apg_gkr_forward_layer", [VP, VP, VP, SZ, SZ, VP])
    d("apg_gkr_forward_layer_dev", [VP, VP, VP, SZ, SZ, VP])
    d("apg_gkr_prove_sumcheck", [VP, VP, VP, VP, SZ, VP, VP, SZ, ROUND_FN, VP, VP, VP, VP])
    d("apg_gkr_prove_sumcheck_dev", [VP, VP, VP, VP, SZ, VP, VP, SZ, ROUND_FN, VP, VP, VP, VP])
    d("apg_transcript_state", [VP, VP])
    d("apg_transcript_set_state", [VP, VP])
    d("apg_fr_to_bytes", [VP, VP])
    d("apg_fr_from_bytes", [VP, VP])
    d("apg_fq_to_bytes", [VP, VP])
    d("apg_fq_from_bytes", [VP, VP])
    d("apg_fr_from_u64", [U64, VP], None)
    d("apg_g1_to_wire", [VP, VP])
    d("apg_g1_from_wire", [VP, VP])
    d("apg_ctx_set_profiling", [VP, I32])
    d("apg_ctx_stats", [VP, C.POINTER(ApgStats)])
    d("apg_ctx_reset_stats", [VP])
    d("apg_synthetic_fill", [U64, I32, SZ, VP])
    d("apg_debug_mul32_batch", [I32, VP, VP, SZ, VP])
    d("apg_debug_poseidon2_batch", [I32, VP, SZ, VP])
    d("apg_pcs_open", [VP, VP, VP, SZ, VP, VP, VP, VP, VP])
    d("apg_pcs_open_dev", [VP, VP, VP, SZ, VP, VP, VP, VP, VP])
    d("apg_pcs_batch_open_stacked", [VP, VP, VP, VP, SZ, SZ, VP, VP, VP, SZ, VP, VP, VP])
    d("apg_pcs_batch_open_stacked_dev", [VP, VP, VP, VP, SZ, SZ, VP, VP, VP, SZ, VP, VP, VP])
    d("apg_debug_fq29_batch", [VP, I32, VP, VP, SZ, VP])
    d("apg_debug_accumulate29", [VP, I32, VP, VP, VP, SZ, VP])
    d("apg_synthetic_fill_dev", [VP, U64, I32, SZ, VP])
    d("apg_engine_create", [I32, I32, SZ, C.POINTER(VP)])
    d("apg_engine_create_lanes", [I32, I32, SZ, I32, I32, C.c_double, C.POINTER(VP)])
    d("apg_engine_destroy", [VP], None)
    d("apg_engine_submit_shard", [VP, VP, VP, SZ, SZ, I32, VP, C.POINTER(C.c_uint64)])
    d("apg_engine_poll", [VP, C.c_uint64])
    d("apg_engine_wait", [VP, C.c_uint64, VP, VP, VP, VP, C.POINTER(ApgJobStats)])
    d("apg_engine_stats", [VP, C.POINTER(ApgEngineInfo)])
    d("apg_engine_set_policy", [VP, I32, I32, C.c_double])
    d("apg_engine_set_preemption", [VP, C.c_double, I32, C.POINTER(C.c_uint64)])
    d("apg_poseidon2_permute_dev", [VP, VP, SZ])
    d("apg_sumcheck_prove_graph_dev", [VP, C.POINTER(VP), SZ, SZ, VP, VP, VP, VP, VP])
    d("apg_transcript_run_dev", [VP, C.POINTER(VP), SZ, VP, SZ, VP, VP, SZ, VP, SZ])
    d("apg_engine_set_profiling", [VP, I32])
    d("apg_engine_kernel_stats", [VP, C.POINTER(ApgStats)])
    d("apg_engine_reset_kernel_stats", [VP])


class ApgJobStats(C.Structure):
    _fields_ = [("queue_ms", C.c_double), ("run_ms", C.c_double), ("total_ms", C.c_double),
                ("worker", C.c_int)]


class ApgEngineInfo(C.Structure):
    _fields_ = [(f, C.c_size_t) for f in ("queued", "running", "completed", "hbm_in_use", "hbm_peak",
                                          "hbm_budget", "arena_bytes", "workers")]


class ApgStats(C.Structure):
    _fields_ = [
        ("accumulate_ms", C.c_double),
        ("accumulate_launches", C.c_uint64),
        ("accumulate_entries", C.c_uint64),
        ("msm_ms", C.c_double),
        ("msm_calls", C.c_uint64),
        ("msm_points", C.c_uint64),
        ("sumcheck_kernel_ms", C.c_double),
        ("sumcheck_launches", C.c_uint64),
        ("sumcheck_bytes", C.c_uint64),
        ("sumcheck_ms", C.c_double),
    ]


def synthetic(seed: int, n: int, kind: str = "uniform") -> np.ndarray:
    """n seeded Montgomery Fr elements: kind "uniform" or "witness" (20% zero,
    10% one, 50% u32, 20% uniform Fr), from ChaCha20 keyed by seed."""
    out = np.zeros((n, 4), dtype=np.uint64)
    _check(lib().apg_synthetic_fill(seed, {"uniform": 0, "witness": 1}[kind], n, _p(out)))
    return out


def _check(rc: int):
    if rc != 0:
        raise ApgError(rc, lib().apg_last_error().decode(errors="replace"))


def _p(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(VP)


def _fe_array(x) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(x, dtype=np.uint64))
    return a.reshape(-1, 4)


# ------------------------------------------------------------------ encodings
def fr_from_u64(x: int) -> np.ndarray:
    out = np.zeros(4, dtype=np.uint64)
    lib().apg_fr_from_u64(x, _p(out))
    return out


def fr_from_bytes(b: bytes) -> np.ndarray:
    src = np.frombuffer(bytes(b), dtype=np.uint8).copy()
    out = np.zeros(4, dtype=np.uint64)
    _check(lib().apg_fr_from_bytes(_p(src), _p(out)))
    return out


def fr_to_bytes(a) -> bytes:
    src = np.ascontiguousarray(np.asarray(a, dtype=np.uint64).reshape(4))
    out = np.zeros(32, dtype=np.uint8)
    _check(lib().apg_fr_to_bytes(_p(src), _p(out)))
    return out.tobytes()


def g1_to_wire(mont: bytes) -> bytes:
    src = np.frombuffer(bytes(mont), dtype=np.uint8).copy()
    out = np.zeros(64, dtype=np.uint8)
    _check(lib().apg_g1_to_wire(_p(src), _p(out)))
    return out.tobytes()


def g1_from_wire(wire: bytes) -> bytes:
    src = np.frombuffer(bytes(wire), dtype=np.uint8).copy()
    out = np.zeros(64, dtype=np.uint8)
    _check(lib().apg_g1_from_wire(_p(src), _p(out)))
    return out.tobytes()


# ------------------------------------------------------------------ context
class Context:
    """One HIP stream and scratch arena on one device (apg_ctx)."""

    def __init__(self, device: int = 0):
        h = VP()
        _check(lib().apg_ctx_create(device, C.byref(h)))
        self.h = h
        self.profiling = False
        _live["ctx"].add(self)

    def close(self):
        if self.h:
            lib().apg_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        _check(lib().apg_ctx_synchronize(self.h))

    def set_profiling(self, on: bool = True):
        """Kernel timing with HIP events (apg_stats); the host clocks of MSMs and
        sumchecks are kept either way."""
        _check(lib().apg_ctx_set_profiling(self.h, 1 if on else 0))
        self.profiling = bool(on)

    def stats(self) -> dict:
        s = ApgStats()
        _check(lib().apg_ctx_stats(self.h, C.byref(s)))
        return {name: getattr(s, name) for name, _ in ApgStats._fields_}

    def reset_stats(self):
        _check(lib().apg_ctx_reset_stats(self.h))


_default_ctx: Optional[Context] = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


class DeviceBuffer:
    """Device memory owned by the library's allocator (apg_dev_alloc)."""

    def __init__(self, ctx: Context, nbytes: int):
        self.ctx = ctx
        self.nbytes = nbytes
        p = VP()
        _check(lib().apg_dev_alloc(ctx.h, nbytes, C.byref(p)))
        self.ptr = p
        _live["buffer"].add(self)

    @classmethod
    def from_array(cls, ctx: Context, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(ctx, a.nbytes)
        _check(lib().apg_dev_upload(ctx.h, b.ptr, _p(a), a.nbytes))
        return b

    def download(self, dtype=np.uint64) -> np.ndarray:
        out = np.zeros(self.nbytes // np.dtype(dtype).itemsize, dtype=dtype)
        _check(lib().apg_dev_download(self.ctx.h, _p(out), self.ptr, self.nbytes))
        return out

    def copy_from(self, src: "DeviceBuffer", nbytes: Optional[int] = None):
        """Device-to-device copy by the library's streaming kernel (apg_dev_copy)."""
        nbytes = min(self.nbytes, src.nbytes) if nbytes is None else nbytes
        if nbytes > min(self.nbytes, src.nbytes):
            raise ValueError("copy_from: nbytes exceeds a buffer")
        _check(lib().apg_dev_copy(self.ctx.h, self.ptr, src.ptr, nbytes))

    def free(self):
        if self.ptr and self.ctx.h:
            _check(lib().apg_dev_free(self.ctx.h, self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ------------------------------------------------------------------ MSM
class MsmError(Exception):
    """crates/curve/src/msm.rs:29-33"""


class LengthMismatch(MsmError):
    def __init__(self, bases: int, scalars: int):
        super().__init__(f"LengthMismatch {{ bases: {bases}, scalars: {scalars} }}")
        self.bases, self.scalars = bases, scalars


def _bases_array(bases) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(bases, dtype=np.uint64).reshape(-1, 8))


def msm(bases, scalars, ctx: Optional[Context] = None) -> bytes:
    """curve::msm::msm: sum_i scalars[i] * bases[i] (bases affine Montgomery
    (n, 8) uint64; scalars Montgomery Fr (n, 4)). Returns the affine result,
    64 Montgomery bytes (all-zero = identity)."""
    b = _bases_array(bases)
    s = _fe_array(scalars) if len(scalars) else np.zeros((0, 4), dtype=np.uint64)
    if b.shape[0] != s.shape[0]:
        raise LengthMismatch(b.shape[0], s.shape[0])
    ctx = ctx or default_context()
    out = np.zeros(64, dtype=np.uint8)
    _check(lib().apg_msm_g1_fr(ctx.h, _p(b), _p(s), s.shape[0], _p(out)))
    return out.tobytes()


def msm_small_u32(bases, scalars, ctx: Optional[Context] = None) -> bytes:
    """curve::msm::msm_small_u32 over u32 scalars."""
    b = _bases_array(bases)
    s = np.ascontiguousarray(np.asarray(scalars, dtype=np.uint32).reshape(-1))
    if b.shape[0] != s.shape[0]:
        raise LengthMismatch(b.shape[0], s.shape[0])
    ctx = ctx or default_context()
    out = np.zeros(64, dtype=np.uint8)
    _check(lib().apg_msm_g1_u32(ctx.h, _p(b), _p(s), s.shape[0], _p(out)))
    return out.tobytes()


def set_window_bits(log2n: int, c: int):
    """Pippenger window width for the size class floor(log2 n) = log2n
    (0 restores the built-in table). Results never depend on it."""
    _check(lib().apg_msm_set_window_bits(log2n, c))


def window_bits(n: int) -> int:
    return int(lib().apg_msm_window_bits(n))


def set_slice_rounds(log2n: int, rounds: int):
    """Accumulation slices per resident thread for the size class floor(log2 n) =
    log2n (0 restores the built-in table). Results never depend on it."""
    _check(lib().apg_msm_set_slice_rounds(log2n, rounds))


def slice_rounds(n: int) -> int:
    return int(lib().apg_msm_slice_rounds(n))


# ------------------------------------------------------------------ SRS + commit
class PcsError(Exception):
    """crates/pcs-verify/src/lib.rs:40-87 (the variants this path raises)."""


class UnsupportedNumVars(PcsError):
    def __init__(self, num_vars: int):
        super().__init__(f"UnsupportedNumVars {{ num_vars: {num_vars} }}")
        self.num_vars = num_vars


class SrsTooSmall(PcsError):
    def __init__(self, needed: int, available: int):
        super().__init__(f"SrsTooSmall {{ needed: {needed}, available: {available} }}")
        self.needed, self.available = needed, available


def check_num_vars(num_vars: int) -> int:
    """crates/pcs-verify/src/lib.rs:97-103: n = 2^(2t), 1 <= t <= 27."""
    if not (2 <= num_vars <= MAX_NUM_VARS) or num_vars % 2:
        raise UnsupportedNumVars(num_vars)
    return 1 << num_vars


class SrsError(Exception):
    """crates/srs/src/lib.rs SrsError: the archive's refusals."""


class SrsIo(SrsError):
    pass


class SrsTruncated(SrsError):
    pass


class SrsBadMagic(SrsError):
    pass


class SrsBadVersion(SrsError):
    def __init__(self, version: int):
        super().__init__(f"bad archive version {version}")
        self.version = version


class SrsBadSection(SrsError):
    pass


class SrsInvalidPoint(SrsError):
    def __init__(self, index: int):
        super().__init__(f"invalid point at index {index}")
        self.index = index


def _srs_error(code: int, detail: int) -> SrsError:
    return {1: lambda: SrsIo("i/o error"), 2: lambda: SrsTruncated("truncated archive"),
            3: lambda: SrsBadMagic("bad magic"), 4: lambda: SrsBadVersion(detail),
            5: lambda: SrsBadSection("archive G1 count is not 2^power"),
            6: lambda: SrsInvalidPoint(detail)}[code]()


def srs_archive_check(path: str):
    """The archive's header rules and its two G2 points, on the host alone:
    returns (power, g2 bytes) or raises the SrsError."""
    pw, code, detail = C.c_uint32(), I32(), U64()
    g2 = np.zeros(256, dtype=np.uint8)
    rc = lib().apg_srs_archive_check(os.fsencode(path), C.byref(pw), _p(g2), C.byref(code), C.byref(detail))
    if rc != 0:
        if code.value:
            raise _srs_error(code.value, detail.value)
        _check(rc)
    return pw.value, g2.tobytes()


def g2_wire_valid(wire: bytes) -> bool:
    """G2Affine::from_bytes's validity of a 128-byte point (on the host)."""
    b = np.frombuffer(bytes(wire), dtype=np.uint8).copy()
    if b.shape[0] != 128:
        raise ValueError("a G2 point is 128 bytes")
    return lib().apg_g2_wire_valid(_p(b)) == 1


def batch_to_affine(jac, ctx: Optional[Context] = None) -> np.ndarray:
    """G1Projective::batch_to_affine on the device: (n, 12) Jacobian limbs to
    (n, 8) affine limbs (all-zero for the identity)."""
    ctx = ctx or default_context()
    j = np.ascontiguousarray(np.asarray(jac, dtype=np.uint64).reshape(-1, 12))
    out = np.zeros((j.shape[0], 8), dtype=np.uint64)
    _check(lib().apg_g1_batch_to_affine(ctx.h, _p(j), j.shape[0], _p(out)))
    return out


class Srs:
    """The powers [tau^i]_1, resident in HBM (crates/srs/src/lib.rs:96)."""

    def __init__(self, handle, ctx: Context):
        self.h = handle
        self.ctx = ctx
        _live["srs"].add(self)

    @classmethod
    def upload(cls, bases, ctx: Optional[Context] = None) -> "Srs":
        ctx = ctx or default_context()
        b = _bases_array(bases)
        h = VP()
        _check(lib().apg_srs_upload(ctx.h, _p(b), b.shape[0], C.byref(h)))
        return cls(h, ctx)

    @classmethod
    def generate(cls, tau, n: int, ctx: Optional[Context] = None) -> "Srs":
        """[tau^i]G for i < n, built on the device."""
        ctx = ctx or default_context()
        t = np.ascontiguousarray(np.asarray(tau, dtype=np.uint64).reshape(4))
        h = VP()
        _check(lib().apg_srs_generate(ctx.h, _p(t), n, C.byref(h)))
        return cls(h, ctx)

    @classmethod
    def load(cls, path: str, ctx: Optional[Context] = None):
        """Srs::load (crates/srs/src/lib.rs:200-240): returns (srs, g2 bytes),
        every G1 power validated on the device; raises the SrsError."""
        ctx = ctx or default_context()
        h, code, detail = VP(), I32(), U64()
        g2 = np.zeros(256, dtype=np.uint8)
        rc = lib().apg_srs_load(ctx.h, os.fsencode(path), C.byref(h), _p(g2), C.byref(code), C.byref(detail))
        if rc != 0:
            if code.value:
                raise _srs_error(code.value, detail.value)
            _check(rc)
        return cls(h, ctx), g2.tobytes()

    @classmethod
    def from_wire(cls, wire: bytes, ctx: Optional[Context] = None) -> "Srs":
        """Canonical 64-byte G1 points decoded and validated on the device
        (crates/srs/src/ptau.rs:307-328); SrsInvalidPoint for a bad one."""
        ctx = ctx or default_context()
        n = len(wire) // 64
        b = np.frombuffer(wire, dtype=np.uint8) if n else np.zeros(64, dtype=np.uint8)
        h, bad = VP(), U64()
        rc = lib().apg_srs_upload_wire(ctx.h, _p(b), n, C.byref(h), C.byref(bad))
        if rc != 0:
            if rc == -2 and bad.value < n:
                raise SrsInvalidPoint(bad.value)
            _check(rc)
        return cls(h, ctx)

    def save(self, path: str, g2: bytes):
        """Srs::save (crates/srs/src/lib.rs:175-192) with the given G2 points."""
        b = np.frombuffer(bytes(g2), dtype=np.uint8).copy()
        if b.shape[0] != 256:
            raise ValueError("the two G2 points are 256 bytes")
        _check(lib().apg_srs_save(self.ctx.h, self.h, _p(b), os.fsencode(path)))

    def __len__(self):
        return lib().apg_srs_len(self.h)

    def g1(self, offset: int = 0, n: Optional[int] = None) -> np.ndarray:
        n = len(self) - offset if n is None else n
        out = np.zeros((n, 8), dtype=np.uint64)
        _check(lib().apg_srs_download(self.ctx.h, self.h, offset, n, _p(out)))
        return out

    def prepare(self, n: int):
        _check(lib().apg_srs_prepare(self.ctx.h, self.h, n))

    def device_bytes(self) -> int:
        return lib().apg_srs_device_bytes(self.h)

    def release(self):
        if self.h:
            lib().apg_srs_release(self.h)
            self.h = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


_SMALL = (np.bool_, np.uint8, np.uint16, np.uint32)


def _column_kind(col: np.ndarray) -> str:
    if col.dtype.type in _SMALL:
        return "u32"
    if col.dtype == np.uint64 and col.ndim == 2 and col.shape[1] == 4:
        return "fr"
    raise TypeError("a column is Fr limbs (n, 4) uint64 or a u1/u8/u16/u32 array")


def kzg_commit(srs: Srs, coeffs, ctx: Optional[Context] = None) -> bytes:
    """srs::kzg::kzg_commit (crates/srs/src/kzg.rs:20-31): one MSM over the
    first n powers, no size rule. Dispatches on the backing like
    pcs::column_msm (crates/pcs/src/lib.rs:280-298). Returns affine Montgomery."""
    return commit_batch(srs, [coeffs], ctx)[0]


def commit(srs: Srs, f, ctx: Optional[Context] = None) -> bytes:
    """pcs::commit (crates/pcs/src/lib.rs:237-247): check_num_vars, then the
    MSM of the evaluation table against the first n powers."""
    n = len(f)
    nv = n.bit_length() - 1
    if n != 1 << nv:
        raise ValueError("a multilinear table holds a power of two entries")
    check_num_vars(nv)
    if len(srs) < n:
        raise SrsTooSmall(n, len(srs))
    return kzg_commit(srs, f, ctx)


def commit_batch(srs: Srs, columns: Sequence, ctx: Optional[Context] = None) -> list:
    """Every column of a shard in one call (prover::commit_all,
    crates/prover/src/lib.rs:444-455): same-kind, same-length columns go to
    the device together. Returns affine Montgomery points, one per column."""
    ctx = ctx or srs.ctx
    cols = [np.asarray(c) for c in columns]
    if not cols:
        return []
    n = len(cols[0])
    if any(len(c) != n for c in cols):
        raise ValueError("commit_batch: columns of one call share a length")
    if len(srs) < n:
        raise SrsTooSmall(n, len(srs))
    kinds = {_column_kind(c) for c in cols}
    k = len(cols)
    out = np.zeros(64 * k, dtype=np.uint8)
    if kinds == {"u32"}:
        s = np.ascontiguousarray(np.stack([c.astype(np.uint32) for c in cols]))
        _check(lib().apg_commit_u32(ctx.h, srs.h, _p(s), n, k, _p(out)))
    elif kinds == {"fr"}:
        s = np.ascontiguousarray(np.stack(cols))
        _check(lib().apg_commit_fr(ctx.h, srs.h, _p(s), n, k, _p(out)))
    else:
        return [kzg_commit(srs, c, ctx) for c in cols]
    return [out[64 * i: 64 * (i + 1)].tobytes() for i in range(k)]


def commit_device(srs: Srs, columns: DeviceBuffer, n: int, k: int, ctx: Optional[Context] = None,
                  offset_elems: int = 0) -> list:
    """k Fr columns of n scalars resident in device memory (column-major),
    committed in one call (apg_commit_fr_dev). Affine Montgomery points."""
    ctx = ctx or srs.ctx
    if len(srs) < n:
        raise SrsTooSmall(n, len(srs))
    if 32 * (offset_elems + n * k) > columns.nbytes:
        raise ValueError("commit_device: columns exceed the buffer")
    out = np.zeros(64 * k, dtype=np.uint8)
    ptr = C.c_void_p(columns.ptr.value + 32 * offset_elems)
    _check(lib().apg_commit_fr_dev(ctx.h, srs.h, ptr, n, k, _p(out)))
    return [out[64 * i: 64 * (i + 1)].tobytes() for i in range(k)]


# ------------------------------------------------------------------ transcript
class DegenerateChallenge(PcsError):
    """PcsError::DegenerateChallenge (crates/pcs-verify/src/lib.rs:40-87)."""


class PointLengthMismatch(PcsError):
    def __init__(self, point: int, num_vars: int):
        super().__init__(f"PointLengthMismatch {{ point: {point}, num_vars: {num_vars} }}")
        self.point, self.num_vars = point, num_vars


MERCURY_PROOF_BYTES = 704
APG_ERR_DEGENERATE = -7


def open(srs: "Srs", f, cm: bytes, u, transcript: "Transcript", ctx: Optional["Context"] = None,
         num_vars: Optional[int] = None):
    """pcs::open (crates/pcs/src/lib.rs:313-492): the Mercury opening of the
    commitment ``cm`` (64-byte wire form, absorbed as passed) at the point
    ``u`` for the multilinear ``f``, advancing ``transcript``. ``f`` is the
    evaluation table (Montgomery Fr, numpy ``(2^num_vars, 4)``) or a
    ``DeviceBuffer`` holding it (then give ``num_vars``). Returns ``(v,
    proof)``: v = fhat(u) as Montgomery limbs and the 704 bytes of
    ``MercuryProof::to_bytes``. Raises the reference's errors."""
    ctx = ctx or srs.ctx
    u = _fe_array(u) if len(u) else np.zeros((0, 4), dtype=np.uint64)
    if isinstance(f, DeviceBuffer):
        if num_vars is None:
            raise ValueError("open: give num_vars with a device table")
        nv = num_vars
    else:
        f = _fe_array(f)
        nv = f.shape[0].bit_length() - 1
        if f.shape[0] != 1 << nv:
            raise ValueError("a multilinear table holds a power of two entries")
    n = check_num_vars(nv)
    if u.shape[0] != nv:
        raise PointLengthMismatch(u.shape[0], nv)
    if len(srs) < n:
        raise SrsTooSmall(n, len(srs))
    cmb = np.frombuffer(bytes(cm), dtype=np.uint8).copy()
    v = np.zeros(4, dtype=np.uint64)
    proof = np.zeros(MERCURY_PROOF_BYTES, dtype=np.uint8)
    if isinstance(f, DeviceBuffer):
        rc = lib().apg_pcs_open_dev(ctx.h, srs.h, f.ptr, nv, _p(cmb), _p(u), transcript.h, _p(v), _p(proof))
    else:
        rc = lib().apg_pcs_open(ctx.h, srs.h, _p(f), nv, _p(cmb), _p(u), transcript.h, _p(v), _p(proof))
    if rc == APG_ERR_DEGENERATE:
        raise DegenerateChallenge("DegenerateChallenge")
    _check(rc)
    return v, proof.tobytes()


pcs_open = open  # the module-level name without shadowing concerns for importers


class EmptyBatch(PcsError):
    """PcsError::EmptyBatch."""


class BatchLengthMismatch(PcsError):
    def __init__(self, commitments: int, paired: int):
        super().__init__(f"BatchLengthMismatch {{ commitments: {commitments}, paired: {paired} }}")
        self.commitments, self.paired = commitments, paired


class MixedColumnSizes(PcsError):
    def __init__(self, expected: int, found: int):
        super().__init__(f"MixedColumnSizes {{ expected: {expected}, found: {found} }}")
        self.expected, self.found = expected, found


def batch_open_stacked(srs: "Srs", stacks, cms: Sequence[bytes], u, r, transcript: "Transcript",
                       ctx: Optional["Context"] = None):
    """pcs::batch_open_stacked (crates/pcs/src/lib.rs:701-789): stacks[i] is
    stack i's columns in slot order (Montgomery Fr tables of one size), opened
    as one Mercury instance at u || r. Returns (vs, proof): one value per stack
    and the 704-byte proof of cm* = sum rho^i cm_i. The reference's refusals,
    in its order."""
    ctx = ctx or srs.ctx
    if not stacks or any(len(s) == 0 for s in stacks):
        raise EmptyBatch("EmptyBatch")
    if len(stacks) != len(cms):
        raise BatchLengthMismatch(len(cms), len(stacks))
    cols = [_fe_array(c) for s in stacks for c in s]
    width = cols[0].shape[0].bit_length() - 1
    for c in cols:
        if c.shape[0] != cols[0].shape[0]:
            raise MixedColumnSizes(width, c.shape[0].bit_length() - 1)
    if cols[0].shape[0] != 1 << width:
        raise ValueError("a multilinear table holds a power of two entries")
    n = check_num_vars(width + len(r))
    u = _fe_array(u) if len(u) else np.zeros((0, 4), dtype=np.uint64)
    if u.shape[0] != width:
        raise PointLengthMismatch(u.shape[0], width)
    slots = 1 << len(r)
    for s in stacks:
        if len(s) > slots:
            raise BatchLengthMismatch(slots, len(s))
    if len(srs) < n:
        raise SrsTooSmall(n, len(srs))
    flat = np.ascontiguousarray(np.concatenate(cols))
    sizes = np.array([len(s) for s in stacks], dtype=np.uint32)
    cmb = np.frombuffer(b"".join(bytes(c) for c in cms), dtype=np.uint8).copy()
    ra = _fe_array(r) if len(r) else np.zeros((1, 4), dtype=np.uint64)
    ua = u if u.shape[0] else np.zeros((1, 4), dtype=np.uint64)
    vs = np.zeros((len(stacks), 4), dtype=np.uint64)
    proof = np.zeros(MERCURY_PROOF_BYTES, dtype=np.uint8)
    rc = lib().apg_pcs_batch_open_stacked(ctx.h, srs.h, _p(flat), _p(sizes), len(stacks), width, _p(cmb), _p(ua),
                                          _p(ra), len(r), transcript.h, _p(vs), _p(proof))
    if rc == APG_ERR_DEGENERATE:
        raise DegenerateChallenge("DegenerateChallenge")
    _check(rc)
    return vs, proof.tobytes()


def batch_open(srs: "Srs", cols, cms: Sequence[bytes], u, transcript: "Transcript", ctx: Optional["Context"] = None):
    """pcs::batch_open (crates/pcs/src/lib.rs:689-698): k same-size columns at
    one point, one proof of cm* = sum rho^i cm_i."""
    return batch_open_stacked(srs, [[c] for c in cols], cms, u, [], transcript, ctx)


class Transcript:
    """The reference's Poseidon2 duplex transcript, native."""

    def __init__(self, _h=None):
        if _h is None:
            _h = VP()
            _check(lib().apg_transcript_new(C.byref(_h)))
        self.h = _h

    def __del__(self):
        try:
            if self.h:
                lib().apg_transcript_free(self.h)
        except Exception:
            pass

    def clone(self) -> "Transcript":
        h = VP()
        _check(lib().apg_transcript_clone(self.h, C.byref(h)))
        return Transcript(h)

    def observe(self, x):
        lib().apg_transcript_observe(self.h, _p(_fe_array(x)))

    def sample(self) -> np.ndarray:
        out = np.zeros(4, dtype=np.uint64)
        lib().apg_transcript_sample(self.h, _p(out))
        return out

    def append_scalar(self, tag: int, x):
        self.append_scalars(tag, [np.asarray(x, dtype=np.uint64).reshape(4)])

    def append_scalars(self, tag: int, xs):
        a = _fe_array(xs) if len(xs) else np.zeros((0, 4), dtype=np.uint64)
        lib().apg_transcript_append_scalars(self.h, tag, _p(a) if a.size else None, a.shape[0])

    def append_bytes(self, tag: int, data: bytes):
        buf = np.frombuffer(bytes(data) + b"\0", dtype=np.uint8).copy()
        lib().apg_transcript_append_bytes(self.h, tag, _p(buf), len(data))

    def challenge_scalar(self, tag: int) -> np.ndarray:
        out = np.zeros(4, dtype=np.uint64)
        lib().apg_transcript_challenge(self.h, tag, _p(out))
        return out

    def append_g1_points(self, tag: int, wires: Sequence[bytes]):
        buf = np.frombuffer(b"".join(wires) + b"\0", dtype=np.uint8).copy()
        lib().apg_transcript_append_g1(self.h, tag, _p(buf), len(wires))

    def state(self) -> bytes:
        out = np.zeros(232, dtype=np.uint8)
        _check(lib().apg_transcript_state(self.h, _p(out)))
        return out.tobytes()

    def set_state(self, state: bytes):
        """The inverse of state(): the 232-byte sponge state, buffers and fill counts."""
        buf = np.frombuffer(bytes(state), dtype=np.uint8).copy()
        if buf.size != 232:
            raise ValueError("a transcript state is 232 bytes")
        _check(lib().apg_transcript_set_state(self.h, _p(buf)))


def poseidon2_permute_device(states, ctx: Optional[Context] = None) -> np.ndarray:
    """The width-3 Poseidon2 permutation (crates/transcript/src/lib.rs:121-146) of
    n states (n x 3 Montgomery Fr elements, uint64 (n, 3, 4)) on the device,
    one state per thread."""
    ctx = ctx or default_context()
    a = np.ascontiguousarray(np.asarray(states, dtype=np.uint64).reshape(-1, 3, 4)).copy()
    if a.shape[0]:
        _check(lib().apg_poseidon2_permute_dev(ctx.h, _p(a), a.shape[0]))
    return a


class TranscriptProgram:
    """Transcript operations recorded for one Transcript and run on the device
    with others (run_transcripts_device): the typed layer (crates/transcript/src/
    lib.rs:306-346) as raw observe and sample steps. sample() and challenge()
    return the index of their output in the run's result."""

    OBSERVE, OBSERVE_U64, SAMPLE = 0, 1, 2

    def __init__(self, transcript: "Transcript"):
        self.transcript = transcript
        self.ops = []        # (op, a, b)
        self.operands = []   # Montgomery elements, uint64 x 4
        self.samples = 0

    def observe(self, x):
        self.ops.append((self.OBSERVE, len(self.operands), 0))
        self.operands.append(np.asarray(x, dtype=np.uint64).reshape(4))

    def observe_u64(self, v: int):
        self.ops.append((self.OBSERVE_U64, v & 0xFFFFFFFF, v >> 32))

    def sample(self) -> int:
        self.ops.append((self.SAMPLE, self.samples, 0))
        self.samples += 1
        return self.samples - 1

    def append_scalars(self, tag: int, xs):
        xs = list(xs)
        self.observe_u64(tag)
        self.observe_u64(len(xs))
        for x in xs:
            self.observe(x)

    def append_bytes(self, tag: int, data: bytes):
        self.observe_u64(tag)
        self.observe_u64(len(data))
        for off in range(0, len(data), 31):
            self.observe(fr_from_bytes(bytes(data[off:off + 31]).ljust(32, b"\0")))

    def challenge(self, tag: int) -> int:
        self.observe_u64(tag)
        return self.sample()


def run_transcripts_device(programs: Sequence[TranscriptProgram], ctx: Optional[Context] = None) -> list:
    """Runs every program on its transcript on the device, one transcript per
    thread (apg_transcript_run_dev); the transcripts take their new states.
    Returns each program's sampled elements, in order."""
    ctx = ctx or default_context()
    if not programs:
        return []
    ops, offs, operands, out_base = [], [0], [], []
    n_out = 0
    for prog in programs:
        base_in = len(operands)
        out_base.append(n_out)
        for op, a, b in prog.ops:
            if op == TranscriptProgram.OBSERVE:
                a += base_in
            elif op == TranscriptProgram.SAMPLE:
                a += n_out
            ops.append((op, a, b, 0))
        operands.extend(prog.operands)
        n_out += prog.samples
        offs.append(len(ops))
    op_arr = np.array(ops, dtype=np.uint32).reshape(-1, 4) if ops else np.zeros((1, 4), dtype=np.uint32)
    off_arr = np.array(offs, dtype=np.uint32)
    in_arr = np.stack(operands).astype(np.uint64) if operands else np.zeros((1, 4), dtype=np.uint64)
    out_arr = np.zeros((max(n_out, 1), 4), dtype=np.uint64)
    hs = (VP * len(programs))(*[p.transcript.h for p in programs])
    _check(lib().apg_transcript_run_dev(ctx.h, hs, len(programs), _p(op_arr), len(ops), _p(off_arr), _p(in_arr),
                                        len(operands), _p(out_arr), n_out))
    return [[out_arr[out_base[i] + j].copy() for j in range(p.samples)] for i, p in enumerate(programs)]


# ------------------------------------------------------------------ sumcheck
@dataclass
class GateTerm:
    """coef * inputs[a] * inputs[b] (* inputs[c]) (crates/sumcheck/src/lib.rs:88-96);
    the third factor is the degree-4 extension."""

    coef: np.ndarray
    a: int
    b: Optional[int] = None
    c: Optional[int] = None


class GateError(Exception):
    pass


class Gate:
    """crates/sumcheck/src/lib.rs:113-151: a formula over declared inputs."""

    def __init__(self, inputs: Sequence[int], terms: Sequence[GateTerm]):
        if not inputs:
            raise GateError("NoInputs")
        for i in range(len(inputs)):
            for j in range(i):
                if inputs[i] == inputs[j]:
                    raise GateError(f"DuplicateInput {{ input: {i} }}")
        for t, term in enumerate(terms):
            for idx in (term.a, term.b, term.c):
                if idx is not None and idx >= len(inputs):
                    raise GateError(f"TermIndexOutOfRange {{ term: {t}, index: {idx} }}")
        self.inputs = list(inputs)
        self.terms = list(terms)

    @property
    def arity(self) -> int:
        return len(self.inputs)

    def degree(self) -> int:
        mx = max([1 + (t.b is not None) + (t.c is not None) for t in self.terms] + [1])
        return max(3, mx + 1)

    def c_terms(self):
        arr = (ApgTerm * max(1, len(self.terms)))()
        for i, t in enumerate(self.terms):
            coef = np.ascontiguousarray(np.asarray(t.coef, dtype=np.uint64).reshape(4)).view(np.uint8)
            for k in range(32):
                arr[i].coef_mont[k] = int(coef[k])
            fs = [t.a, -1 if t.b is None else t.b, -1 if t.c is None else t.c]
            for k in range(3):
                arr[i].factor[k] = fs[k]
        return arr

    @classmethod
    def product(cls, k: int) -> "Gate":
        """x_0 * x_1 * ... * x_(k-1) with coefficient one (k <= 3)."""
        one = fr_from_u64(1)
        fs = list(range(k)) + [None] * (3 - k)
        return cls(list(range(k)), [GateTerm(one, fs[0], fs[1], fs[2])])


@dataclass
class SumcheckProof:
    """crates/sumcheck/src/lib.rs:179-186: rounds[i] ascending coefficients."""

    rounds: np.ndarray  # (n, d + 1, 4)
    final_evals: np.ndarray  # (arity, 4)


def _tables_host(tables):
    arrs = [np.ascontiguousarray(np.asarray(t, dtype=np.uint64).reshape(-1, 4)) for t in tables]
    ptrs = (VP * len(arrs))(*[a.ctypes.data for a in arrs])
    return arrs, ptrs


def prove_sumcheck(eq_point, gate: Gate, tables, transcript: Transcript, ctx: Optional[Context] = None):
    """gkr::prove_sumcheck with a gate summand: returns (rounds, point, final_evals).
    Each round is absorbed as SUMCHECK_ROUND and followed by a
    SUMCHECK_CHALLENGE on ``transcript`` (crates/gkr/src/lib.rs:886-897)."""
    ctx = ctx or default_context()
    arrs, ptrs = _tables_host(tables)
    n = int(arrs[0].shape[0]).bit_length() - 1
    if any(a.shape[0] != 1 << n for a in arrs):
        raise ValueError("prove_sumcheck: tables share one power-of-two length")
    eqp = _fe_array(eq_point) if n else np.zeros((1, 4), dtype=np.uint64)
    if n and eqp.shape[0] != n:
        raise ValueError("prove_sumcheck: the eq point has one coordinate per variable")
    d = gate.degree()
    rounds = np.zeros((max(n, 1), d + 1, 4), dtype=np.uint64)
    point = np.zeros((max(n, 1), 4), dtype=np.uint64)
    finals = np.zeros((len(arrs), 4), dtype=np.uint64)
    terms = gate.c_terms()
    cb = C.cast(lib().apg_transcript_round, ROUND_FN)
    _check(lib().apg_sumcheck_prove(ctx.h, terms, len(gate.terms), ptrs, len(arrs), n, _p(eqp), cb,
                                    transcript.h, _p(rounds), _p(point), _p(finals)))
    return rounds[:n], point[:n], finals


BACKINGS = {"fr": 0, "u1": 1, "u8": 2, "u16": 3, "u32": 4}


def backing_of(column: np.ndarray) -> str:
    """The PolyBacking a host column maps to: bool -> u1, uint8/16/32 -> u8/16/32,
    (n, 4) uint64 Montgomery -> fr (crates/poly/src/lib.rs:46-58)."""
    if column.dtype == np.bool_:
        return "u1"
    if column.dtype == np.uint8:
        return "u8"
    if column.dtype == np.uint16:
        return "u16"
    if column.dtype == np.uint32:
        return "u32"
    if column.dtype == np.uint64 and column.ndim == 2 and column.shape[1] == 4:
        return "fr"
    raise ValueError(f"no table backing for dtype {column.dtype} shape {column.shape}")


def upload_backed(ctx: Context, column: np.ndarray) -> tuple:
    """A host column to device memory at its own width: (DeviceBuffer, backing).
    Bits are packed into u64 limbs, entry i at bit i % 64 of limb i / 64."""
    kind = backing_of(column)
    if kind == "u1":
        n = column.shape[0]
        bits = np.zeros(((n + 63) // 64) * 64, dtype=np.uint8)
        bits[:n] = column.astype(np.uint8)
        data = np.packbits(bits, bitorder="little").view(np.uint64)
    else:
        data = np.ascontiguousarray(column)
    return DeviceBuffer.from_array(ctx, data), kind


def prove_sumcheck_device(eq_point, gate: Gate, tables, transcript: Transcript, ctx: Optional[Context] = None):
    """prove_sumcheck over device tables given as (DeviceBuffer, backing)
    pairs, narrow backings lifted on the device (apg_sumcheck_prove_dev_lifted);
    every table holds 2^n entries. Returns (rounds, point, final_evals)."""
    ctx = ctx or default_context()
    n = len(eq_point)
    ptrs = (C.c_void_p * len(tables))(*[b.ptr.value for b, _ in tables])
    kinds = np.array([BACKINGS[k] for _, k in tables], dtype=np.int32)
    eqp = _fe_array(eq_point) if n else np.zeros((1, 4), dtype=np.uint64)
    d = gate.degree()
    rounds = np.zeros((max(n, 1), d + 1, 4), dtype=np.uint64)
    point = np.zeros((max(n, 1), 4), dtype=np.uint64)
    finals = np.zeros((len(tables), 4), dtype=np.uint64)
    cb = C.cast(lib().apg_transcript_round, ROUND_FN)
    _check(lib().apg_sumcheck_prove_dev_lifted(ctx.h, gate.c_terms(), len(gate.terms), ptrs, _p(kinds), len(tables),
                                               n, _p(eqp), cb, transcript.h, _p(rounds), _p(point), _p(finals)))
    return rounds[:n], point[:n], finals


def prove_sumcheck_graph(eq_point, tables, transcript: Transcript, ctx: Optional[Context] = None):
    """The product sumcheck eq * t_0 * ... * t_(k-1) (k <= 3) over DeviceBuffers
    with every round and transcript step on the device, one captured graph per
    proof (apg_sumcheck_prove_graph_dev). Returns (rounds, point, final_evals);
    the transcript advances as the host-driven path's would."""
    ctx = ctx or default_context()
    n, k = len(eq_point), len(tables)
    ptrs = (C.c_void_p * k)(*[b.ptr.value for b in tables])
    eqp = _fe_array(eq_point)
    rounds = np.zeros((n, max(4, k + 2), 4), dtype=np.uint64)
    point = np.zeros((n, 4), dtype=np.uint64)
    finals = np.zeros((k, 4), dtype=np.uint64)
    _check(lib().apg_sumcheck_prove_graph_dev(ctx.h, ptrs, k, n, _p(eqp), transcript.h, _p(rounds), _p(point),
                                              _p(finals)))
    return rounds, point, finals


def witness_digest(columns) -> np.ndarray:
    """sumcheck::witness_digest (crates/sumcheck/src/lib.rs:228-256): the
    Poseidon2 sponge over every cell (host, sequential). Returns one Fr."""
    if len(columns) == 0:
        raise ValueError("witness_digest: a witness has at least one column")
    arrs, ptrs = _tables_host(columns)
    n = int(arrs[0].shape[0]).bit_length() - 1
    for k, a in enumerate(arrs):
        if a.shape[0] != 1 << n:
            raise ValueError(f"witness_digest: column {k} has {a.shape[0]} cells, column 0 has {1 << n}")
    out = np.zeros(4, dtype=np.uint64)
    _check(lib().apg_witness_digest(ptrs, len(arrs), n, _p(out)))
    return out


def absorb_witness_digest(t: Transcript, digest) -> None:
    """crates/sumcheck/src/lib.rs:262: append_scalars(WITNESS_DIGEST, [digest])."""
    t.append_scalars(TAG_WITNESS_DIGEST, np.asarray(digest, dtype=np.uint64).reshape(1, 4))


def prove_zerocheck(gate: Gate, polys, transcript: Transcript, ctx: Optional[Context] = None) -> SumcheckProof:
    """sumcheck::prove_zerocheck: the caller has already absorbed the witness
    digest; draws r, proves, absorbs final_evals."""
    ctx = ctx or default_context()
    if len(polys) != gate.arity:
        raise ValueError(f"prove_zerocheck: {len(polys)} columns for a gate with {gate.arity} inputs")
    arrs, ptrs = _tables_host(polys)
    n = int(arrs[0].shape[0]).bit_length() - 1
    for k, a in enumerate(arrs):
        if a.shape[0] != 1 << n:
            raise ValueError(f"prove_zerocheck: column {k} has {a.shape[0]} entries, column 0 has {1 << n}")
    d = gate.degree()
    rounds = np.zeros((max(n, 1), d + 1, 4), dtype=np.uint64)
    finals = np.zeros((len(arrs), 4), dtype=np.uint64)
    _check(lib().apg_zerocheck_prove(ctx.h, gate.c_terms(), len(gate.terms), ptrs, len(arrs), n, transcript.h,
                                     _p(rounds), _p(finals)))
    return SumcheckProof(rounds[:n], finals)


@dataclass
class ShardProof:
    commitments: list  # wire G1, one per column
    rounds: np.ndarray  # (num_vars, degree + 1, 4)
    final_evals: np.ndarray  # (degree - 1, 4)


def prove_shard(srs: Srs, columns: DeviceBuffer, num_vars: int, k: int, degree: int,
                transcript: Transcript, ctx: Optional[Context] = None) -> ShardProof:
    """Commit k device-resident columns of 2^num_vars Fr, then the batched
    sumcheck over their rho-combinations (apg_prove_shard_dev)."""
    ctx = ctx or srs.ctx
    cms = np.zeros(64 * k, dtype=np.uint8)
    rounds = np.zeros((num_vars, degree + 1, 4), dtype=np.uint64)
    finals = np.zeros((degree - 1, 4), dtype=np.uint64)
    _check(lib().apg_prove_shard_dev(ctx.h, srs.h, columns.ptr, num_vars, k, degree, transcript.h,
                                     _p(cms), _p(rounds), _p(finals)))
    return ShardProof([cms[64 * i: 64 * (i + 1)].tobytes() for i in range(k)], rounds, finals)


def synthetic_device(ctx: Context, seed: int, n: int, kind: str = "witness", out: Optional[DeviceBuffer] = None,
                     offset_elems: int = 0) -> DeviceBuffer:
    """n counter-mode ChaCha20 elements (the values `synthetic` returns) written
    straight into device memory."""
    buf = out or DeviceBuffer(ctx, 32 * n)
    ptr = C.c_void_p(buf.ptr.value + 32 * offset_elems)
    _check(lib().apg_synthetic_fill_dev(ctx.h, seed, {"uniform": 0, "witness": 1}[kind], n, ptr))
    return buf


def imul_peak(ctx: Optional[Context] = None, iters: int = 4096, reps: int = 5) -> dict:
    """The measured issue rate of v_mad_u64_u32 (apg_debug_imul_peak)."""
    ctx = ctx or default_context()
    ms, rate = C.c_double(), C.c_double()
    _check(lib().apg_debug_imul_peak(ctx.h, iters, reps, C.byref(ms), C.byref(rate)))
    return {"best_ms": ms.value, "imul_per_s": rate.value}


# ------------------------------------------------------------------ engine
class Engine:
    """Continuous batching of shard jobs on one device (apg_engine): `workers`
    concurrent HIP streams, admission against an HBM budget in arrival order,
    shortest first or longest first (set_policy)."""

    def __init__(self, device: int = 0, workers: int = 4, hbm_budget: int = 200 << 30, lane_workers: int = 0,
                 lane_cus: int = 0, small_scalars: float = float(1 << 18)):
        """lane_workers > 0: those workers are a fast lane for jobs of at most
        small_scalars rows x columns (shortest-first), on CUs [0, lane_cus)
        apart from the rest (apg_engine_create_lanes). Engine streams run
        concurrently only with enough hardware queues: GPU_MAX_HW_QUEUES >=
        2 x workers in the environment before HIP starts."""
        h = VP()
        if lane_workers > 0:
            _check(lib().apg_engine_create_lanes(device, workers, hbm_budget, lane_workers, lane_cus,
                                                 float(small_scalars), C.byref(h)))
        else:
            _check(lib().apg_engine_create(device, workers, hbm_budget, C.byref(h)))
        self.h = h
        self._jobs = {}
        self._preempt = (0.0, "size")
        _live["engine"].add(self)

    def close(self):
        if self.h:
            lib().apg_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit_shard(self, srs: Srs, columns: DeviceBuffer, num_vars: int, k: int, degree: int,
                     seed: Optional[Transcript] = None, column_offset_elems: int = 0) -> int:
        jid = C.c_uint64()
        ptr = C.c_void_p(columns.ptr.value + 32 * column_offset_elems)
        _check(lib().apg_engine_submit_shard(self.h, srs.h, ptr, num_vars, k, degree,
                                             seed.h if seed is not None else None, C.byref(jid)))
        self._jobs[jid.value] = (num_vars, k, degree)
        return jid.value

    def done(self, job: int) -> bool:
        """Whether the job has finished; ApgError for an id never issued or
        already waited for."""
        rc = lib().apg_engine_poll(self.h, job)
        _check(min(rc, 0))
        return rc == 1

    def wait(self, job: int):
        """-> (ShardProof, final Transcript, stats dict); ApgError for an id
        never issued or already waited for."""
        if job not in self._jobs:
            raise ApgError(-2, "engine: unknown or already collected job id")
        num_vars, k, degree = self._jobs.pop(job)
        cms = np.zeros(64 * k, dtype=np.uint8)
        rounds = np.zeros((num_vars, degree + 1, 4), dtype=np.uint64)
        finals = np.zeros((degree - 1, 4), dtype=np.uint64)
        t = Transcript()
        st = ApgJobStats()
        _check(lib().apg_engine_wait(self.h, job, _p(cms), _p(rounds), _p(finals), t.h, C.byref(st)))
        proof = ShardProof([cms[64 * i: 64 * (i + 1)].tobytes() for i in range(k)], rounds, finals)
        return proof, t, {"queue_ms": st.queue_ms, "run_ms": st.run_ms, "total_ms": st.total_ms,
                          "worker": st.worker}

    def stats(self) -> dict:
        s = ApgEngineInfo()
        _check(lib().apg_engine_stats(self.h, C.byref(s)))
        return {name: getattr(s, name) for name, _ in ApgEngineInfo._fields_}

    POLICIES = {"fifo": 0, "shortest": 1, "longest": 2}

    def set_policy(self, policy: str = "fifo", reserve: int = 0, small_scalars: float = 0.0):
        """Queue order: "fifo", "shortest" or "longest" (rows x columns); with reserve > 0,
        jobs above small_scalars hold at most workers - reserve workers."""
        if policy not in self.POLICIES:
            raise ValueError(f"unknown policy {policy!r}")
        _check(lib().apg_engine_set_policy(self.h, self.POLICIES[policy], reserve, float(small_scalars)))

    PREEMPT = {"size": 0, "urgency": 1}

    def set_preemption(self, ratio: float = 0.0, mode: str = "size") -> int:
        """Phase-boundary preemption: with ratio >= 1 a running job pauses at its next
        phase boundary while a job it gives way to runs beside it: one ratio times
        smaller (rows x columns, mode "size"), or one whose urgency (latency so far
        plus estimated remaining time, over its estimated time alone) exceeds its own
        times ratio (mode "urgency"). 0 turns it off. Returns the pauses so far."""
        if mode not in self.PREEMPT:
            raise ValueError(f"unknown preemption mode {mode!r}")
        pauses = C.c_uint64(0)
        _check(lib().apg_engine_set_preemption(self.h, float(ratio), self.PREEMPT[mode], C.byref(pauses)))
        self._preempt = (float(ratio), mode)
        return int(pauses.value)

    def pauses(self) -> int:
        """Pauses of jobs at phase boundaries so far (the preemption setting is kept)."""
        return self.set_preemption(*self._preempt)

    def set_profiling(self, on: bool = True):
        _check(lib().apg_engine_set_profiling(self.h, 1 if on else 0))

    def kernel_stats(self) -> dict:
        s = ApgStats()
        _check(lib().apg_engine_kernel_stats(self.h, C.byref(s)))
        return {name: getattr(s, name) for name, _ in ApgStats._fields_}

    def reset_kernel_stats(self):
        _check(lib().apg_engine_reset_kernel_stats(self.h))
