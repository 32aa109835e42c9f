"""CPU-side checks of the C ABI: the library loads, exports every symbol that
include/hipmps.h declares, its host-side gate table matches the oracle, and argument
errors come back as codes (no GPU needed)."""
import ctypes
import os
import re

import numpy as np
import pytest

from oracle import oracle as O
from tnqvm_amd import _capi as C
from tnqvm_amd import mps

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "hipmps.h")).read()
    return sorted(set(re.findall(r"\b(hipmps_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = C.lib()
    declared = header_functions()
    assert len(declared) >= 25
    for name in declared:
        assert hasattr(L, name), name
    assert sorted(C.EXPORTS) == declared


@pytest.mark.parametrize("name,params", [
    ("I", ()), ("H", ()), ("X", ()), ("Y", ()), ("Z", ()), ("S", ()), ("Sdg", ()), ("T", ()), ("Tdg", ()),
    ("Rx", (0.37,)), ("Ry", (-1.2,)), ("Rz", (2.5,)), ("U", (0.1, 2.2, -0.9)),
    ("CNOT", ()), ("CX", ()), ("CZ", ()), ("CY", ()), ("CH", ()), ("CRZ", (0.4,)), ("CPhase", (-0.8,)),
    ("Swap", ()), ("iSwap", ()), ("fSim", (1.1, 0.3))])
def test_gate_table_matches_oracle(name, params):
    assert np.array_equal(mps.gate_matrix(name, params), O.gate_matrix(name, params))


def test_unknown_gate_is_an_error():
    with pytest.raises(C.HipMpsError):
        mps.gate_matrix("NotAGate")
    out = np.zeros(32)
    rc = C.lib().hipmps_gate_matrix(b"Rx", None, 0, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert rc == -1  # Rx without its angle


def test_null_arguments_return_codes():
    L = C.lib()
    assert L.hipmps_engine_create(None, None) == -1
    assert L.hipmps_state_create(None, 4, None, None) == -1
    assert L.hipmps_flush(None) == -1
    assert L.hipmps_engine_destroy(None) == 0
    assert b"null" in L.hipmps_last_error() or len(L.hipmps_last_error()) > 0
