"""Runs the C++ operator-layer test driver (tests/cpp/test_operators.cpp) on
the GPU: the reference's gtests for this path, through the C++ classes."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(__file__), "cpp", "test_operators")

CASES = [
    "SplitTransposeTester.cycling_prime_test",
    "CoherentBeamformerTester.representative_noise_test",
    "IncoherentBeamformerTester.ib_representative_noise_test",
    "WeightsManagerTester.test_zero_value",
    "WeightsManagerTester.test_real_value",
    "DoubleDeviceBuffer.semantics",
    "CoherentBeamformer.error_paths_throw",
    "DelayManagerTester.test_updates",
    "PipelineTester.simple_run_test",
    "Pipeline.outputs_match_operators_16ant",
    "Pipeline.outputs_match_operators_64ant_fused",
    "Pipeline.outputs_match_operators_64of128_taftp",
    "TransposeToDada.beams_match_oracle",
    "FengToDada.reorders_in_place",
]


@pytest.mark.parametrize("case", CASES)
def test_cpp_operator(case):
    assert os.path.exists(EXE), "C++ test driver not built (run __graft_entry__.build())"
    res = subprocess.run([EXE, case], capture_output=True, text=True, timeout=110)
    print(res.stdout[-3000:], res.stderr[-2000:])
    assert res.returncode == 0, res.stdout[-3000:]
    assert "[       OK ] " + case in res.stdout
