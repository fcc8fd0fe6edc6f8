"""C++ gpuNUFFT:: API on the GPU: runs tests/cpp/cpp_api_check (built by
__graft_entry__.build()) and requires every check in it to pass.

The program checks the GpuArray overloads against the host overloads,
loadPrecomputedGpuNUFFTOperator against the factory-built operator, the
operator flavour per factory flags, stop-stage shapes, the adjoint identity
with and without sensitivities, and exceptions for invalid arguments."""

import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "cpp_api_check")


def test_cpp_api_check():
    assert os.path.exists(BIN), "cpp_api_check is not built; run __graft_entry__.build()"
    print("running", BIN, flush=True)
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout[-4000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0 and "CPP_API_CHECK OK" in r.stdout
