"""Host-only cases of the C++ test driver (no device work): DADA header
parsing, the ring buffer's write/read/end-of-data contract within a process
and across processes (a writer process filling two observations that a
DadaInputStream reads with a new read client each, and the one-writer lock),
and the DelayManager's error on a missing shared-memory segment."""
import os
import subprocess

import pytest

EXE = os.path.join(os.path.dirname(__file__), "cpp", "test_operators")


@pytest.mark.parametrize("case", ["Header.get_set", "DadaRing.write_read_eod", "DadaRing.cross_process_two_observations",
                                  "DadaRing.writer_lock_across_processes", "DelayManagerTester.test_bad_keys"])
def test_cpp_host_case(case):
    if not os.path.exists(EXE):
        pytest.skip("C++ test driver not built (run __graft_entry__.build())")
    res = subprocess.run([EXE, case], capture_output=True, text=True, timeout=60)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "[       OK ] " + case in res.stdout
