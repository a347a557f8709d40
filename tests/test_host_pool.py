"""The engine's host planning pool (lcurve_re_amd/csrc/lc_pool.h) through
tests/cxx/pool_check.cpp: every item of 26 000 jobs of random sizes runs
exactly once (jobs started while the workers sleep, spin, or are still
leaving the previous job; items that sleep), every exception thrown by an
item reaches the caller, and two threads using the pool at once both
complete (the one that finds it busy runs inline).  CPU only."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "cxx", "_build", "pool_check")


@pytest.mark.skipif(not os.path.exists(EXE), reason="tests/cxx not built (run __graft_entry__.build())")
def test_host_pool_runs_every_item_once():
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "pool ok" in r.stdout
