"""Host-only check of the ensemble launch-table cache's eviction rules
(mom5_amd/csrc/table_cache.h): compiled with g++, no GPU involved."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_pinned_tables_are_never_evicted(tmp_path):
    exe = tmp_path / "table_cache_check"
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-I",
                    os.path.join(ROOT, "mom5_amd", "csrc"),
                    os.path.join(ROOT, "tests", "table_cache_check.cpp"), "-o", str(exe)],
                   check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.strip() == "OK", r.stdout + r.stderr
