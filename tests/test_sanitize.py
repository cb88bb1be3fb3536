"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer: the
oracle (which checks every GPU result) and the serving loop's arena
allocator, built with gcc/g++ on the CPU and run on small problems."""
import os
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g"]


def _run(cmd, exe, marker):
    subprocess.run(cmd, check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert marker in r.stdout


def test_oracle_under_sanitizers(tmp_path):
    exe = tmp_path / "oracle_san"
    _run(["gcc", "-O1", "-std=c11", "-Wall", *SAN, "-I", os.path.join(REPO, "oracle"),
          os.path.join(REPO, "tests", "cpp", "test_oracle_sanitize.c"),
          os.path.join(REPO, "oracle", "tcv_oracle.c"), "-o", str(exe), "-lm", "-lpthread"],
         exe, "oracle sanitize ok")


def test_arena_under_sanitizers(tmp_path):
    exe = tmp_path / "arena_san"
    _run(["g++", "-O1", "-std=c++17", "-Wall", *SAN,
          "-I", os.path.join(REPO, "askap-benchmarks_amd", "csrc"),
          os.path.join(REPO, "tests", "cpp", "test_arena.cpp"), "-o", str(exe)],
         exe, "arena ok")
