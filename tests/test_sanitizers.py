"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer.

* The oracle's C restatement: oracle/Makefile's `asan` target builds the same source with both
  sanitizers; the oracle's own test files then run in a child process that loads that build
  (CPC_ORACLE_LIB) with the ASan runtime preloaded into the Python interpreter.
* The corr_pc program's reader and planner (cpc_main.cpp, the code that parses .inp and catalog
  files): corr_pc_amd/build.py builds it with both sanitizers, and the program's CPU tests
  (tests/test_program.py, `--plan`, which stops before the GPU) run against that build (CPC_PROG).

Any out-of-bounds access, use after free or undefined behaviour aborts the child, and its report
fails the test.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
SAN_LIB = os.path.join(ORACLE, "_asan", "libcpc_oracle.so")


def _asan_runtime():
    try:
        p = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True, check=True)
    except (OSError, subprocess.CalledProcessError):
        return None
    path = p.stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


def test_sanitizer_runtime_reports_a_planted_overflow(tmp_path):
    """The sanitizer flags of the oracle's build do catch an out-of-bounds write on this host."""
    if _asan_runtime() is None:
        pytest.skip("gcc has no AddressSanitizer runtime here")
    src = tmp_path / "oob.c"
    src.write_text("#include <stdlib.h>\nint main(int c, char **v) { int *a = malloc(4 * sizeof(int));"
                   " a[c + 3] = 1; int r = a[0]; free(a); return r; }\n")
    exe = tmp_path / "oob"
    subprocess.run(["gcc", "-O1", "-g", "-fsanitize=address,undefined", str(src), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, env=dict(os.environ, ASAN_OPTIONS="detect_leaks=0"))
    assert r.returncode != 0 and "heap-buffer-overflow" in r.stderr


def test_oracle_under_address_and_ub_sanitizers():
    rt = _asan_runtime()
    if rt is None:
        pytest.skip("gcc has no AddressSanitizer runtime here")
    subprocess.run(["make", "-s", "-C", ORACLE, "asan"], check=True)
    env = dict(os.environ)
    env.update(LD_PRELOAD=rt, CPC_ORACLE_LIB=SAN_LIB, OMP_NUM_THREADS="4",
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    # the child loads the sanitized build, not the default one
    probe = subprocess.run([sys.executable, "-c", "import sys; sys.path.insert(0, %r); from oracle import oracle; "
                            "print(oracle.load()._name)" % ROOT], capture_output=True, text=True, env=env, cwd=ROOT,
                           timeout=300)
    assert probe.returncode == 0, probe.stderr[-2000:]
    assert probe.stdout.strip() == SAN_LIB
    files = sorted(os.path.join("tests", f) for f in os.listdir(os.path.join(ROOT, "tests"))
                   if f.startswith("test_oracle") and f.endswith(".py"))
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-m", "not gpu"] + files,
                       capture_output=True, text=True, env=env, cwd=ROOT, timeout=900)
    out = r.stdout + r.stderr
    assert "AddressSanitizer" not in out and "runtime error" not in out, out[-4000:]
    assert r.returncode == 0, out[-4000:]
    print(out.strip().splitlines()[-1])


def test_program_reader_under_address_and_ub_sanitizers():
    if _asan_runtime() is None:
        pytest.skip("gcc has no AddressSanitizer runtime here")
    from corr_pc_amd.build import LIB, build_program_sanitized
    if not os.path.exists(LIB):
        pytest.skip("the library is not built")
    prog = build_program_sanitized()
    env = dict(os.environ, CPC_PROG=prog, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-m", "not gpu",
                        os.path.join("tests", "test_program.py")], capture_output=True, text=True, env=env, cwd=ROOT,
                       timeout=900)
    out = r.stdout + r.stderr
    assert "AddressSanitizer" not in out and "runtime error" not in out, out[-4000:]
    assert r.returncode == 0, out[-4000:]
    print(out.strip().splitlines()[-1])
