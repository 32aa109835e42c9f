"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only): the hip-mps visitor
layer through the CPU cases of the C++ visitor tests, and the CPU oracle through a driver that runs
its state vector, SVD split, MPS (with and without truncation), amplitudes, slices and both
samplers.  Built by tnqvm_amd.build.build_sanitized (g++ -fsanitize=address,undefined)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "bin")


def _bin(name):
    path = os.path.join(BIN, name)
    if not os.path.exists(path):
        from tnqvm_amd import build as B
        B.build_sanitized()
    return path


def _run(args, leaks):
    env = dict(os.environ, ASAN_OPTIONS=f"detect_leaks={int(leaks)}:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run(args, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, env=env, timeout=300)
    print(r.stdout[-3000:])
    assert r.returncode == 0, r.stdout[-4000:]
    assert "ERROR: AddressSanitizer" not in r.stdout and "runtime error" not in r.stdout
    return r.stdout


def test_oracle_under_sanitizers():
    out = _run([_bin("oracle_san")], leaks=True)
    assert "0 failures" in out


def test_visitor_cpu_cases_under_sanitizers():
    # static singletons of the service registry live until exit: leak checking off for this binary
    out = _run([_bin("visitor_tests_san"), "--cpu"], leaks=False)
    assert " 0 failures" in out
