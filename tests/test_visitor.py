"""The C++ hip-mps visitor through the TNQVM accelerator API (tests/cpp/visitor_tests.cpp).

CPU cases: XASM front end, lnn-transform routing, option handling.
GPU cases: the exatn-mps gtest cases (gates, routing, sampling, seed determinism, norm,
closed forms, amplitudes / slices, VQE mode, accelerator state) against hip-mps, and the
exatn-pmps gtest cases (ExaTnPmpsTester.cpp) against hip-pmps.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "bin", "visitor_tests")


def _binary():
    if not os.path.exists(BIN):
        from tnqvm_amd import build as B
        B.build_all()
    return BIN


def _run(flag, timeout):
    r = subprocess.run([_binary(), flag], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=timeout)
    print(r.stdout)
    assert r.returncode == 0, r.stdout[-4000:]
    assert " 0 failures" in r.stdout


def test_visitor_cpu_cases():
    _run("--cpu", 120)


@pytest.mark.gpu
def test_visitor_gpu_cases():
    _run("--gpu", 600)
