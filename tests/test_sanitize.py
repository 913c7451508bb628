"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer, as the reference builds
its tests with NWGRAPH_SANITIZE (CMakeLists.txt:84-88): the file readers and writers of
nwgraph_amd/csrc/io.cpp driven by tests/cpp/io_check.cpp (valid files, round trips,
malformed and oversized headers, every truncation, seeded corruptions), and the oracle's own
tests rerun against a sanitizer build of the oracle.  CPU only."""
import os
import subprocess
import sys

import pytest

from conftest import GOLDEN, ROOT

SAN = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-fno-sanitize-recover=all"]


def _runtime(name):
    p = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True).stdout.strip()
    return p if os.path.isabs(p) and os.path.exists(p) else None


def test_io_readers_under_sanitizers(tmp_path):
    exe = tmp_path / "io_check"
    build = subprocess.run(["g++", "-std=c++17", *SAN, "-I", os.path.join(ROOT, "include"),
                            os.path.join(ROOT, "tests", "cpp", "io_check.cpp"),
                            os.path.join(ROOT, "nwgraph_amd", "csrc", "io.cpp"), "-o", str(exe)],
                           capture_output=True, text=True, timeout=300)
    assert build.returncode == 0, build.stderr[-4000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    run = subprocess.run([str(exe), GOLDEN, str(tmp_path)], capture_output=True, text=True, timeout=600, env=env)
    assert run.returncode == 0, (run.stdout[-2000:], run.stderr[-4000:])
    assert "all reader checks passed" in run.stdout


def test_oracle_tests_under_sanitizers(tmp_path):
    asan, ubsan = _runtime("libasan.so"), _runtime("libubsan.so")
    if not (asan and ubsan):
        pytest.skip("gcc's sanitizer runtimes are not installed")
    build = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize", f"SANDIR={tmp_path}"],
                           capture_output=True, text=True, timeout=300)
    assert build.returncode == 0, build.stderr[-4000:]
    env = dict(os.environ, LD_PRELOAD=f"{asan}:{ubsan}", ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1",
               NWG_ORACLE_LIB=str(tmp_path / "libnwg_oracle_san.so"), OMP_NUM_THREADS="4")
    run = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                          os.path.join(ROOT, "tests", "test_oracle.py")],
                         capture_output=True, text=True, timeout=900, env=env, cwd=ROOT)
    assert run.returncode == 0, (run.stdout[-3000:], run.stderr[-3000:])
    assert " passed" in run.stdout
