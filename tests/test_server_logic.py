"""CPU tests of the serving loop's host logic (no GPU): the resident arena's
first-fit allocator and the request size classes, compiled from
fishfry_amd/csrc/ff_arena.hpp with g++ and checked against a model."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_arena_and_size_classes(tmp_path):
    exe = tmp_path / "arena_test"
    subprocess.run(["g++", "-O1", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "fishfry_amd", "csrc"),
                    os.path.join(ROOT, "tests", "cpp", "arena_test.cpp"), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "arena ok" in r.stdout
