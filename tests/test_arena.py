"""The serving loop's HBM-budget allocator (host logic), compiled with g++ and
run on the CPU: admission depends on it never over-committing the budget."""
import os
import subprocess

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_arena_allocator(tmp_path):
    exe = tmp_path / "test_arena"
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-Werror",
                    "-I", os.path.join(REPO, "askap-benchmarks_amd", "csrc"),
                    os.path.join(REPO, "tests", "cpp", "test_arena.cpp"), "-o", str(exe)],
                   check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "arena ok" in r.stdout
