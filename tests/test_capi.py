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


# Closed forms that pin both gate tables (the library's and the oracle's) independently of
# either: Pauli algebra, rotations as exp(-i theta P / 2), U as a ZYZ Euler product, controlled
# gates as |0><0| (x) I + |1><1| (x) U with U from the pinned 1-qubit gates (control = first
# operand = most significant bit, tnqvm/base/Gates.hpp:237-246), and the two-qubit exchange
# gates as exponentials of XX + YY.  A sign, conjugation or operand-order slip fails here.
def _expm(a):
    from scipy.linalg import expm
    return expm(a)


@pytest.mark.parametrize("table", ["library", "oracle"])
def test_gate_tables_closed_forms(table):
    gm = (lambda n, p=(): mps.gate_matrix(n, p)) if table == "library" else (lambda n, p=(): O.gate_matrix(n, p))
    I2 = np.eye(2)
    X = np.array([[0, 1], [1, 0]], dtype=complex)
    Z = np.diag([1.0, -1.0]).astype(complex)
    Y = 1j * X @ Z
    P0, P1 = np.diag([1.0, 0.0]), np.diag([0.0, 1.0])

    def ctrl(u):
        return np.kron(P0, I2) + np.kron(P1, u)

    tol = 1e-15
    assert np.allclose(gm("I"), I2, atol=0)
    assert np.allclose(gm("X"), X, atol=0)
    assert np.allclose(gm("Z"), Z, atol=0)
    assert np.allclose(gm("Y"), Y, atol=0)  # Y = i X Z
    assert np.allclose(gm("H"), (X + Z) / np.sqrt(2), atol=tol)
    assert np.allclose(gm("T"), np.diag([1, np.exp(0.25j * np.pi)]), atol=tol)
    assert np.allclose(gm("Tdg"), gm("T").conj().T, atol=tol)
    assert np.allclose(gm("S"), np.diag([1, 1j]), atol=tol)
    assert np.allclose(gm("Sdg"), gm("S").conj().T, atol=tol)
    rng = np.random.default_rng(17)
    for th, ph, la in rng.uniform(-np.pi, np.pi, (6, 3)):
        for name, pauli in (("Rx", X), ("Ry", Y), ("Rz", Z)):
            assert np.allclose(gm(name, (th,)), _expm(-0.5j * th * pauli), atol=1e-14), name
        # U(theta, phi, lambda) = e^{i (phi + lambda)/2} Rz(phi) Ry(theta) Rz(lambda)
        euler = np.exp(0.5j * (ph + la)) * (_expm(-0.5j * ph * Z) @ _expm(-0.5j * th * Y) @ _expm(-0.5j * la * Z))
        assert np.allclose(gm("U", (th, ph, la)), euler, atol=1e-14)
        assert np.allclose(gm("CRZ", (th,)), ctrl(_expm(-0.5j * th * Z)), atol=1e-14)
        assert np.allclose(gm("CPhase", (th,)), ctrl(np.diag([1, np.exp(1j * th)])), atol=1e-14)
        # fSim(theta, phi) = exp(-i theta (XX + YY) / 2) . CPhase(-phi)
        xy = np.kron(X, X) + np.kron(Y, Y)
        assert np.allclose(gm("fSim", (th, ph)), _expm(-0.5j * th * xy) @ ctrl(np.diag([1, np.exp(-1j * ph)])),
                           atol=1e-14)
    assert np.allclose(gm("CNOT"), ctrl(X), atol=0)
    assert np.allclose(gm("CX"), ctrl(X), atol=0)
    assert np.allclose(gm("CY"), ctrl(Y), atol=0)
    assert np.allclose(gm("CZ"), ctrl(Z), atol=0)
    assert np.allclose(gm("CH"), ctrl((X + Z) / np.sqrt(2)), atol=tol)
    # CH = (I x Ry(pi/4)) CZ (I x Ry(-pi/4)): H is Z rotated by pi/4 about y
    ry = lambda a: _expm(-0.5j * a * Y)  # noqa: E731
    assert np.allclose(gm("CH"), np.kron(I2, ry(np.pi / 4)) @ ctrl(Z) @ np.kron(I2, ry(-np.pi / 4)), atol=1e-14)
    swap = np.eye(4)[[0, 2, 1, 3]]
    assert np.allclose(gm("Swap"), swap, atol=0)
    # iSwap = exp(+i pi/4 (XX + YY))
    assert np.allclose(gm("iSwap"), _expm(0.25j * np.pi * (np.kron(X, X) + np.kron(Y, Y))), atol=1e-14)


# The ctypes mirrors of the header's structs, checked field by field against what a C compiler
# makes of include/hipmps.h: size, and each field's offset and size.  A field added to one side
# only, or reordered, fails here before any call could read the wrong bytes.
_STRUCTS = [(C.EngineOpts, "hipmps_engine_opts"), (C.TruncOpts, "hipmps_trunc_opts"), (C.Gate, "hipmps_gate"),
            (C.Tuning, "hipmps_tuning"), (C.Stats, "hipmps_stats_t"), (C.Circuit, "hipmps_circuit"),
            (C.Result, "hipmps_result"), (C.Profile, "hipmps_profile_t")]


def test_ctypes_structs_match_header(tmp_path):
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    lines = ['#include <stdio.h>', '#include "hipmps.h"', "int main(void) {"]
    for cls, cname in _STRUCTS:
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in cls._fields_:
            lines.append(f'  printf("{cname} {f} %zu %zu\\n", offsetof({cname}, {f}), sizeof((({cname}*)0)->{f}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run([cc, "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {}
    for ln in out:
        if ln:
            parts = ln.split()
            got[(parts[0], parts[1])] = tuple(int(x) for x in parts[2:])
    for cls, cname in _STRUCTS:
        assert got[(cname, "size")] == (ctypes.sizeof(cls),), cname
        for f, _ in cls._fields_:
            d = getattr(cls, f)
            assert got[(cname, f)] == (d.offset, d.size), (cname, f)
