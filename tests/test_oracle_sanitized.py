"""The oracle's own tests against its AddressSanitizer + UndefinedBehavior-
Sanitizer build (`make -C oracle sanitize` -> liboracle_san.so), run in a
child interpreter with the sanitizer runtimes preloaded: an out-of-bounds
access, a use after free or undefined arithmetic anywhere in the C
restatement fails the run (undefined behaviour is not recoverable in that
build)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")


def _runtime(name):
    p = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True).stdout.strip()
    return p if os.path.isabs(p) and os.path.exists(p) else None


def test_oracle_tests_pass_under_asan_ubsan():
    asan, ubsan = _runtime("libasan.so"), _runtime("libubsan.so")
    if not asan or not ubsan:
        pytest.skip("gcc's sanitizer runtimes are not installed")
    subprocess.run(["make", "-C", ORACLE, "sanitize"], check=True, capture_output=True)
    env = dict(os.environ)
    env.update(LD_PRELOAD=f"{asan}:{ubsan}", ASAN_OPTIONS="detect_leaks=0:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1", FELOG_ORACLE_LIB=os.path.join(ORACLE, "liboracle_san.so"))
    probe = subprocess.run([sys.executable, "-c",
                            "import sys; sys.path.insert(0, %r); import orc; orc.lib();"
                            "print(open('/proc/self/maps').read().count('liboracle_san.so') > 0)" % ORACLE],
                           env=env, capture_output=True, text=True, timeout=120)
    assert probe.returncode == 0 and probe.stdout.strip() == "True", probe.stderr[-2000:]
    r = subprocess.run([sys.executable, "-m", "pytest", os.path.join(ROOT, "tests", "test_oracle.py"), "-x", "-q",
                        "-m", "not gpu", "-p", "no:cacheprovider"], env=env, capture_output=True, text=True,
                       timeout=900, cwd=ROOT)
    tail = (r.stdout + r.stderr)[-4000:]
    assert r.returncode == 0, tail
    assert "ERROR: AddressSanitizer" not in tail and "runtime error" not in tail, tail
