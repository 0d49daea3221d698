"""AddressSanitizer + UndefinedBehaviorSanitizer runs on the CPU (no GPU):

  - the host side of the library (libereb_hip_san.so, built by
    erebus.jl_amd/build.py --sanitize): slot-table discovery, the save_state
    field table, the null-context and argument-validation paths of every
    entry point and of ereb_ctx_create (tests/san_host_checks.py), in a
    child process that preloads clang's ASan runtime;
  - the oracle (oracle/_build_san, `make -C oracle san`: the checker, the
    OpenMP CPU-baseline harness and the dual-number restatement): the oracle
    suites run against it in a child process that preloads gcc's libasan.
Both builds abort on the first report, so a memory error or undefined
behaviour fails the test with the sanitizer's report."""
from __future__ import annotations

import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "erebus.jl_amd")
sys.path.insert(0, PKG)

import build as hipbuild  # noqa: E402  (erebus.jl_amd/build.py)

ORACLE_SUITES = ["tests/test_oracle.py", "tests/test_oracle_hydration.py",
                 "tests/test_oracle_backtrace.py", "tests/test_oracle_replenish.py",
                 "tests/test_oracle_dual.py", "tests/test_oracle_bench.py"]


def _env(preload: str, **extra) -> dict:
    env = dict(os.environ)
    env.update(LD_PRELOAD=preload, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1",
               PYTHONPATH=os.pathsep.join([PKG, os.path.join(ROOT, "tests")]), **extra)
    return env


def _report(r) -> str:
    return f"status {r.returncode}\n{r.stdout[-3000:]}\n{r.stderr[-6000:]}"


def test_library_host_code_under_sanitizers():
    if not os.path.exists(hipbuild.LIB_SAN):
        hipbuild.build_sanitized()
    r = subprocess.run([sys.executable, os.path.join(ROOT, "tests", "san_host_checks.py"),
                        hipbuild.LIB_SAN], cwd=ROOT, capture_output=True, text=True,
                       timeout=600, env=_env(hipbuild.sanitizer_runtime()))
    assert r.returncode == 0 and "host checks passed" in r.stdout, _report(r)


def test_oracle_under_sanitizers():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "san"], check=True)
    libasan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True,
                             text=True, check=True).stdout.strip()
    assert os.path.isabs(libasan), libasan
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                        "-p", "no:xdist", "-m", "not gpu", *ORACLE_SUITES], cwd=ROOT,
                       capture_output=True, text=True, timeout=1200,
                       env=_env(libasan, EREB_ORACLE_BUILD="_build_san"))
    assert r.returncode == 0, _report(r)
