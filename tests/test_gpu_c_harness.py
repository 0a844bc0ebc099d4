"""The C harness (tests/harness/bv_sequence.c, built by build()) drives the
library through its C ABI in the call sequences SLEPc's BV and Mat code
issue through a registered type (gramschmidt slot passes with the DGKS loop
on the host, q == NULL coefficient-buffer dot/mult, active-column scale,
MultInPlace with a wider host Q, BVMult/BVCopy/BVResize/BVDot/BVMatMult,
MatMult, BVMatArnoldi) and checks each result with plain C loops."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "harness", "bv_sequence")


def test_c_harness_slepc_call_sequences():
    assert os.path.exists(BIN), "build the harness first (__graft_entry__.build)"
    p = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(p.stdout)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "PASSED: 0 failure(s)" in p.stdout
    assert p.stdout.count("\nok  ") + p.stdout.startswith("ok  ") >= 20
