import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running CPU reference check")


TWIN_SRC = os.path.join(ROOT, "tests", "host_twin.hip")
TWIN_LIB = os.path.join(ROOT, "tests", "_twin", "libfew_twin.so")
CSRC = os.path.join(ROOT, "fastemriwaveforms_amd", "csrc")


def build_twin() -> str:
    """Compile the host twin of the kernel sources (hipcc, host code exercised on the CPU)."""
    srcs = [TWIN_SRC] + [os.path.join(CSRC, f) for f in os.listdir(CSRC)]
    if os.path.exists(TWIN_LIB) and os.path.getmtime(TWIN_LIB) >= max(map(os.path.getmtime, srcs)):
        return TWIN_LIB
    os.makedirs(os.path.dirname(TWIN_LIB), exist_ok=True)
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = TWIN_LIB + ".tmp"
    subprocess.run([hipcc, "-O2", "-fPIC", "-std=c++17", "--offload-arch=gfx950",
                    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                    "-I", CSRC, "-I", os.path.join(ROOT, "include"), "-shared", TWIN_SRC, "-o", tmp],
                   check=True, capture_output=True)
    os.replace(tmp, TWIN_LIB)
    return TWIN_LIB


@pytest.fixture(scope="session")
def twin():
    import ctypes
    return ctypes.CDLL(build_twin())


@pytest.fixture(scope="session")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    from fastemriwaveforms_amd import _lib
    _lib.lib()  # raises if libfew_rocm.so is missing: GPU tests never run without it
    return torch.device("cuda", 0)
