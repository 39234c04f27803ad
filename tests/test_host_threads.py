"""The host plan builders' thread pool (tgh_forest.hpp for_block_chunks),
compiled for the CPU with g++: an exception on a worker reaches the caller
after every thread has been joined, instead of terminating the process."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def test_a_worker_exception_reaches_the_caller(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ is not installed")
    exe = str(tmp_path / "chunks_throw")
    subprocess.run([gxx, "-std=c++17", "-O1", "-pthread", "-o", exe, os.path.join(HERE, "native", "chunks_throw.cpp")],
                   check=True)
    res = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert res.returncode == 0, res.stdout + res.stderr
    assert res.stdout.startswith("ok")
