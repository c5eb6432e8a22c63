"""CPU-only checks of the native library: it loads without a GPU, exports every
symbol include/apogee_gpu.h declares, and its host-side pieces (the Poseidon2
transcript, the encodings, the interface's error behaviour) match the
reference's vectors. No compute call needs a GPU here."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT, golden_lines
from test_oracle_poly_transcript import replay_cases


@pytest.fixture(scope="session")
def apg():
    import importlib

    return importlib.import_module("apogee-vm_amd")


def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "apogee_gpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(apg_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(apg):
    syms = _declared_symbols()
    assert len(syms) >= 40
    lib = ctypes.CDLL(apg.LIB_PATH)
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_native_transcript_replays_reference_cases(apg, orc):
    class Adapter:
        def __init__(self):
            self.t = apg.Transcript()

        def observe(self, x):
            self.t.observe(x)

        def sample(self):
            return self.t.sample()

        def append_scalars(self, tag, xs):
            self.t.append_scalars(tag, np.asarray(xs).reshape(-1, 4))

        def append_bytes(self, tag, b):
            self.t.append_bytes(tag, b)

        def challenge(self, tag):
            return self.t.challenge_scalar(tag)

    assert replay_cases(orc, Adapter) >= 10


def test_native_transcript_state_matches_oracle_after_g1(apg, orc):
    recs = list(golden_lines("g1_absorb_kats.txt", {"single"}))
    wires = [bytes.fromhex(r[1]) for r in recs]
    a, b = apg.Transcript(), orc.Transcript()
    a.append_g1_points(3, wires)
    b.append_g1(3, wires)
    assert np.array_equal(a.challenge_scalar(5), b.challenge(5))


def test_encodings(apg, orc):
    rng = orc.SplitMix(3)
    for _ in range(20):
        x = rng.next_fr()
        assert apg.fr_to_bytes(x) == orc.to_bytes(orc.FR, x)
        assert np.array_equal(apg.fr_from_bytes(orc.to_bytes(orc.FR, x)), x)
    p = (0x30644E72E131A029B85045B68181585D2833E84879B9709143E1F593F0000001).to_bytes(32, "little")
    with pytest.raises(apg.ApgError):
        apg.fr_from_bytes(p)
    assert np.array_equal(apg.fr_from_u64(12345), orc.from_u64(orc.FR, 12345))
    for rec in golden_lines("g1_kats_subset.txt", {"g1_ops"}):
        w = bytes.fromhex(rec[1])
        m = apg.g1_from_wire(w)
        assert apg.g1_to_wire(m) == w
        assert np.array_equal(np.frombuffer(m, dtype=np.uint64), orc.affine_from_wire(w))
    for rec in golden_lines("g1_kats_subset.txt", {"g1_reject"}):
        with pytest.raises(apg.ApgError):
            apg.g1_from_wire(bytes.fromhex(rec[1]))
    assert apg.g1_to_wire(bytes(64)) == bytes(64)


def test_interface_errors(apg):
    one = apg.fr_from_u64(1)
    with pytest.raises(apg.GateError):
        apg.Gate([], [])
    with pytest.raises(apg.GateError):
        apg.Gate([1, 1], [])
    with pytest.raises(apg.GateError):
        apg.Gate([1, 2], [apg.GateTerm(one, 0, 2)])
    for nv in (0, 1, 3, 55, 56):
        with pytest.raises(apg.UnsupportedNumVars):
            apg.check_num_vars(nv)
    assert apg.check_num_vars(4) == 16
    with pytest.raises(apg.LengthMismatch):
        apg.msm(np.zeros((3, 8), dtype=np.uint64), np.zeros((2, 4), dtype=np.uint64))
    assert apg.Gate.product(2).degree() == 3
    assert apg.Gate.product(3).degree() == 4
    assert apg.Gate([0], [apg.GateTerm(one, 0)]).degree() == 3
